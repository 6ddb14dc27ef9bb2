/*
 * qnb.h -- C ABI of the MI355X nonbonded engine for Q's spherical-boundary MD.
 *
 * The engine replaces the nonbonded half of Qdyn6's potential energy: the
 * pair-list builders, the nonbonded energy/gradient loops, the per-state Q-atom
 * FEP/EVB energies, the local-reaction-field (LRF) expansion and the SCAAS
 * water-sphere restraints.  Each entry point below names the reference routine
 * it stands in for (paths relative to the reference's src/).
 *
 * Conventions
 *  - Plain C types only; every pointer is borrowed for the duration of the call.
 *  - Atom, charge-group, Q-atom and type indices are 0-based.  The one
 *    exception is iqatom[], which keeps the reference meaning: 0 = not a
 *    Q-atom, k > 0 = Q-atom number k (1-based), as src/globals.f90:65.
 *  - Coordinates and gradients crossing the ABI are the reference layout:
 *    natom x {x,y,z} float64 (TYPE(qr_vec), src/sizes.f90:109-111).  The
 *    gradient is d = dE/dx (the force is -d), src/potene.f90:109.
 *  - Charges are already scaled by sqrt(coulomb constant), as prep_sim does
 *    (src/simprep.f90:3714-3721).
 *  - Multi-dimensional tables keep the reference's Fortran (column-major)
 *    order, so the Fortran shim can pass its arrays unchanged; the layout of
 *    each one is spelled out next to it.
 *  - Every function returns 0 on success and a nonzero QNB_E* code otherwise;
 *    qnb_last_error() gives the message.  The Fortran side turns a nonzero code
 *    into die(msg) (src/qalloc.f90:609), keeping the reference's semantics.
 *  - A handle is used from one host thread at a time.  It owns its device
 *    memory and HIP stream.
 */
#ifndef QNB_H
#define QNB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QNB_ABI_VERSION 6

enum qnb_status {
  QNB_OK = 0,
  QNB_EINVAL = 1,     /* bad argument or inconsistent system description */
  QNB_EDEVICE = 2,    /* HIP runtime error */
  QNB_ENOMEM = 3,     /* allocation failed */
  QNB_ENONFINITE = 4, /* non-finite energy (reference: NaN check, src/md.f90:1861) */
  QNB_EOVERFLOW = 5,  /* a list exceeded its capacity after regrowth */
  QNB_EUNSUPPORTED = 6
};

enum qnb_precision {
  QNB_PREC_FP64 = 0,  /* plain: FP64 throughout */
  QNB_PREC_MIXED = 1  /* FP32 pair arithmetic, FP64 accumulation */
};

enum qnb_vdw_rule { QNB_VDW_GEOMETRIC = 1, QNB_VDW_ARITHMETIC = 2 };

#define QNB_MAX_NBR_RANGE 25 /* src/topo.f90:37 */
#define QNB_MAX_STATES 8

/* Everything the engine needs from Qdyn6's set-up (topology, FEP file,
 * prep_sim, precompute_interactions).  Sizes first. */
typedef struct qnb_system {
  int32_t natom;       /* all atoms (solute + solvent)                  */
  int32_t nat_solute;  /* solute atoms come first: 0 .. nat_solute-1   */
  int32_t nwat;        /* solvent molecules (each one charge group)     */
  int32_t solv_atom;   /* atoms per solvent molecule (3 for TIP3P/SPC)  */
  int32_t ncgp;        /* charge groups; solvent groups follow solute   */
  int32_t ncgp_solute;
  int32_t nqat;        /* Q-atoms                                        */
  int32_t nstates;     /* FEP/EVB states (1..QNB_MAX_STATES)             */
  int32_t ntypes;      /* atom types (natyps)                            */
  int32_t ivdw_rule;   /* qnb_vdw_rule, src/topo.f90:185                 */
  int32_t solvent_spc; /* 1: SPC-like solvent, only atom 0 carries LJ
                          (nonbond_ww_spc / nonbond_qw_spc paths,
                          src/potene.f90:347-356); 0: general loops     */
  int32_t nexlong;     /* long (> QNB_MAX_NBR_RANGE) exclusions          */
  int32_t n14long;     /* long 1-4 pairs                                 */
  int32_t any_atom_cutoff; /* 1: the topology's iuse_switch_atom = 0: a charge-
                          group pair is inside a cut-off when any atom pair
                          is (nbpplis2, nbpwlis2, nbqplis2 and their _lrf
                          forms, src/nonbondene.f90:796-825); 0 (default):
                          switch atoms (iuse_switch_atom = 1)            */

  /* per atom [natom] */
  const double *crg;     /* scaled charge                                  */
  const int32_t *iac;    /* atom type, 0-based                             */
  const int32_t *iqatom; /* 0 or 1-based Q-atom number                     */
  const uint8_t *excl;   /* 1 = excluded atom (outside the sphere)         */
  const uint8_t *heavy;  /* 1 = heavy atom (SCAAS radial acts on heavy
                            solvent atoms, src/nonbondene.f90:6510-6540)   */

  /* charge groups [ncgp]: atom ranges into cgpatom[], switch atom */
  const int32_t *cgp_first;   /* index into cgpatom                        */
  const int32_t *cgp_last;    /* index into cgpatom (inclusive)            */
  const int32_t *cgp_iswitch; /* atom index                                */
  const int32_t *cgpatom;     /* [cgp_last[ncgp-1]+1] atom indices          */

  /* atom types (iaclib(t)%avdw(k), src/topo.f90:79-83): avdw/bvdw
   * [ntypes][3], element (type t, LJ code k) at [t*3 + k], with code k = 0
   * normal, 1 type-2, 2 the 1-4 parameters.  ljcod [ntypes][ntypes] holds the
   * 1-based code (1 or 2) of a type pair at [ti + tj*ntypes]
   * (src/globals.f90:136). */
  const double *avdw;
  const double *bvdw;
  const int32_t *ljcod;
  double el14_scale;

  /* solute exclusions, the topology's own encoding (src/topo.f90:200-204):
   * listex/list14 [nat_solute][QNB_MAX_NBR_RANGE]: element (k-1, i) at
   * [i*25 + k-1] is 1 when pair (i, i+k) is excluded / is a 1-4 pair;
   * listexlong/list14long [n][2] atom pairs at [2*m], [2*m+1]. */
  const uint8_t *listex;
  const uint8_t *list14;
  const int32_t *listexlong;
  const int32_t *list14long;

  /* solute-solvent parameters: pw_* [solv_atom][nat_solute], element
   * (solute atom i, solvent atom j) at [i + j*nat_solute]
   * (pw_precomp(i-pw_low, j), src/simprep.f90:3069-3108, offset removed). */
  const double *pw_A, *pw_B, *pw_el;
  /* solvent-solvent: ww_* [solv_atom][solv_atom], (la, ka) at
   * [la + ka*solv_atom] (ww_precomp, src/simprep.f90:3296-3317). */
  const double *ww_A, *ww_B, *ww_el;

  /* Q-atoms */
  const int32_t *iqseq; /* [nqat] atom index of each Q-atom                */
  /* qconn3 [nat_solute]: 1 when the solute atom is bonded or angled
   * (qconn <= 3) to any Q-atom in any state, i.e. any(qconn(:,i,:) <= 3)
   * (src/nonbondene.f90:3717); such atoms go on the Q-"bonded neighbour"
   * list instead of the Q-solute list. */
  const uint8_t *qconn3;
  /* qp_* [nstates][nqat][nat_solute]: (i, iq, s) at [i + nat_solute*(iq +
   * nqat*s)] (qp_precomp, src/simprep.f90:3112-3148, 3420-3461) */
  const double *qp_A, *qp_B, *qp_el, *qp_sc;
  const uint8_t *qp_set;
  /* qq_* [nstates][nqat][nqat]: (iq, jq, s) at [iq + nqat*(jq + nqat*s)]
   * (qq_precomp, src/simprep.f90:3152-3268) */
  const double *qq_A, *qq_B, *qq_el, *qq_sc;
  const uint8_t *qq_set, *qq_soft;
  /* qw_* [nstates][solv_atom][nqat]: (iq, j, s) at [iq + nqat*(j +
   * solv_atom*s)] (qw_precomp, src/simprep.f90:3272-3292) */
  const double *qw_A, *qw_B, *qw_el, *qw_sc;

  /* nonbonded pairs inside each solvent molecule (nonbnd_solv_int,
   * prep_sim_precompute_solvent src/simprep.f90:4004-4138; evaluated by
   * nonbond_solvent_internal src/nonbondene.f90:6082-6154 for every solvent
   * molecule, into the solvent-solvent energies): n_solv_int pairs,
   * solv_int_ij [n][2] the two atoms' 0-based offsets in the molecule,
   * solv_int_par [n][3] vdW A, vdW B and the charge product.  Qdyn6 builds
   * the list only when the topology has solvent torsions (0 pairs
   * otherwise); pairs listed twice are evaluated twice, as there. */
  int32_t n_solv_int;
  const int32_t *solv_int_ij;
  const double *solv_int_par;
} qnb_system;

/* Run parameters: cutoffs (Angstrom, not squared), sphere, LRF, SCAAS. */
typedef struct qnb_params {
  double Rcpp, Rcpw, Rcww, Rcq, RcLRF; /* src/globals.f90:132            */
  int32_t use_LRF;
  int32_t precision;                   /* qnb_precision                  */
  /* bit 0: with use_LRF = 0, reproduce nbpwlist's distance test exactly as
   * the reference executes it: its charge-group loop reuses the switch-atom
   * variable `ia`, so after the first water inside the cut-off the test
   * point becomes atom cgp_last+1 (src/nonbondene.f90:2686-2691). */
  int32_t ref_quirks;
  double xpcent[3];                    /* sphere centre (Q-list origin)  */
  double xwcent[3];                    /* water-sphere centre            */
  double rexcl_o;                      /* exclusion radius (Q lists are
                                          static when Rcq > rexcl_o)     */
  double lambda[QNB_MAX_STATES];       /* EQ(s)%lambda                   */
  /* SCAAS radial restraint (src/nonbondene.f90:6466-6544) */
  double rwat, fk_wsphere, Dwmz, awmz, Boltz;
  /* SCAAS polarisation restraint (src/nonbondene.f90:6548-6746) */
  int32_t wpol_restr;
  int32_t nwpolr_shell;                /* <= QNB_MAX_SHELLS              */
  double fkwpol;
  double itdis;                        /* theta_corr update cadence      */
  double wshell_rout[8], wshell_dr[8], wshell_cstb[8];
  double wshell_theta_corr[8];         /* initial (restart) values       */
} qnb_params;

#define QNB_MAX_SHELLS 8

/* E%pp, E%pw, E%ww, E%LRF and the two SCAAS restraint terms
 * (TYPE(ENERGIES), src/nrgy.f90:48-54). */
typedef struct qnb_energies {
  double pp_el, pp_vdw;
  double pw_el, pw_vdw;
  double ww_el, ww_vdw;
  double lrf;
  double solvent_radial, water_pol;
} qnb_energies;

/* Per-state Q-atom energies EQ(s)%qq, %qp, %qw (src/nrgy.f90:64-71). */
typedef struct qnb_qenergies {
  double qq_el, qq_vdw;
  double qp_el, qp_vdw;
  double qw_el, qw_vdw;
} qnb_qenergies;

typedef struct qnb_handle qnb_handle;

/* Which list qnb_get_pairs exports (src/globals.f90:414-437). */
enum qnb_list {
  QNB_LIST_PP = 0,  /* nbpp  (i, j) solute-solute atom pairs              */
  QNB_LIST_PW = 1,  /* nbpw  (i, j) solute atom - solvent atom            */
  QNB_LIST_WW = 2,  /* nbww  as molecule pairs (iw, jw), 0-based waters   */
  QNB_LIST_QP = 3,  /* nbqp  (iq, j) 0-based Q-atom number, solute atom   */
  QNB_LIST_QW = 4,  /* nbqw  as (iq, iw) Q-atom number, 0-based water     */
  QNB_LIST_QQ = 5,  /* nbqq  (iq, jq) of state 0                          */
  QNB_LIST_QQP = 6  /* nbqqp (iq, j) of state 0                           */
};

int32_t qnb_abi_version(void);
const char *qnb_last_error(void);
/* Number of visible devices (0 when no GPU). */
int32_t qnb_device_count(void);

/* Copies the system to device memory and prepares static lists (Q-Q and
 * Q-bonded-neighbour lists: make_nbqqlist + nbqqlist,
 * src/nonbondene.f90:729-745, 3231-3282).  Call once after Qdyn6's
 * precompute_interactions (src/qdyn.f90:153). */
int32_t qnb_create(const qnb_system *sys, const qnb_params *par, qnb_handle **out);
int32_t qnb_destroy(qnb_handle *h);

/* make_pair_lists (src/nonbondene.f90:749-837), spherical boundary,
 * switch-atom lists: charge-group centres, solute-solute, solute-water,
 * water-water, Q-solute and Q-water lists, and the LRF expansion of every
 * charge-group pair beyond the cut-off when use_LRF.  x: natom*3 (host). */
int32_t qnb_make_pair_lists(qnb_handle *h, const double *x);

/* pot_energy_nonbonds (src/potene.f90:320-380) for in_md = .true.:
 * nonbond_ww(_spc), nonbond_qw(_spc), nonbond_pw, nonbond_pp, nonbond_qp
 * and lrf_taylor.  Adds the gradient into d (natom*3, host) and writes
 * E (pp, pw, ww, lrf) and EQ[nstates] (qp, qw). */
int32_t qnb_nonbonded(qnb_handle *h, const double *x, double *d, qnb_energies *E, qnb_qenergies *EQ);

/* nonbond_qq + nonbond_qqp (src/nonbondene.f90:5013-5153): adds into d and
 * writes EQ[s].qq and adds the Q-bonded-neighbour terms to EQ[s].qp. */
int32_t qnb_qq(qnb_handle *h, const double *x, double *d, qnb_qenergies *EQ);

/* restrain_solvent + watpol (src/nonbondene.f90:6466-6746) for one MD step:
 * Tfree is the previous step's free temperature (src/md.f90:1803); istep
 * drives the theta_corr update cadence.  Adds into d, writes
 * E->solvent_radial and E->water_pol. */
int32_t qnb_restraints(qnb_handle *h, const double *x, double *d, double Tfree, int32_t istep, qnb_energies *E);

/* Term selection for qnb_eval_terms: one bit per reference routine. */
enum qnb_term {
  QNB_T_WW = 1 << 0,  /* nonbond_ww_spc / nonbond_ww   src/nonbondene.f90:5886, 5509 */
  QNB_T_QW = 1 << 1,  /* nonbond_qw_spc / nonbond_qw   :5658, 5339 */
  QNB_T_PW = 1 << 2,  /* nonbond_pw                    :4856 */
  QNB_T_PP = 1 << 3,  /* nonbond_pp                    :4694 */
  QNB_T_QP = 1 << 4,  /* nonbond_qp                    :5157 */
  QNB_T_LRF = 1 << 5, /* lrf_taylor                    :507  */
  QNB_T_QQ = 1 << 6,  /* nonbond_qq                    :5013 */
  QNB_T_QQP = 1 << 7, /* nonbond_qqp                   :5085 */
  QNB_T_RAD = 1 << 8, /* restrain_solvent              :6466 */
  QNB_T_POL = 1 << 9,  /* watpol                       :6548 */
  QNB_T_SINT = 1 << 10 /* nonbond_solvent_internal     :6082; its energies go
                          to ww_el / ww_vdw as there (added to nonbond_ww's
                          when QNB_T_WW is selected too) */
};
#define QNB_T_NONBONDED (QNB_T_WW | QNB_T_QW | QNB_T_PW | QNB_T_PP | QNB_T_QP | QNB_T_LRF | QNB_T_SINT)

/* Evaluates the selected terms (mask of qnb_term) at x, adds their gradient
 * into d and writes their energies; unselected fields are left untouched.
 * Tfree/istep are used by the restraint terms only. */
int32_t qnb_eval_terms(qnb_handle *h, uint32_t mask, const double *x, double *d, double Tfree, int32_t istep,
                       qnb_energies *E, qnb_qenergies *EQ);

/* Exports the current list `which` as n (a, b) int32 pairs into ab[2*cap];
 * *n receives the length (also when it exceeds cap, then nothing is copied). */
int32_t qnb_get_pairs(qnb_handle *h, int32_t which, int32_t *ab, int64_t cap, int64_t *n);

/* LRF table after the last list update: [ncgp][43] float64 in the order
 * centre(3), phi0, phi1(3), phi2(3x3), phi3(9x3) of LRF_TYPE
 * (src/globals.f90:503-509). */
int32_t qnb_get_lrf(qnb_handle *h, double *lrf);

/* Polarisation restraint state that survives a restart (theta_corr,
 * src/qalloc.f90:840-871) plus the running averages: [3*nwpolr_shell] as
 * theta_corr[], avtheta[], avn_insh[]. */
int32_t qnb_get_shell_state(qnb_handle *h, double *state);
int32_t qnb_set_shell_state(qnb_handle *h, const double *state);

/* ------------------------------------------------------------------------
 * Q-atom bonded terms per state: the FEP file's changing bonds, angles,
 * torsions and impropers with their Morse couplings (TYPE(QBOND_TYPE) ...
 * src/qatom.f90:72-122, as get_fep leaves them).  Atom and term numbers are
 * 0-based; cod[t*nstates + s] is the 1-based row of the term's library used in
 * state s, 0 = term absent in s.  Angles in radians, as the reference holds
 * them after reading (src/qatom.f90:1112, 1312, 1506). */
typedef struct qnb_qbonded {
  int32_t nbond, nangle, ntorsion, nimproper;
  int32_t nbond_lib, nangle_lib, ntorsion_lib, nimproper_lib;
  const int32_t *bond_ij, *bond_cod;   /* [nbond][2], [nbond][nstates]               */
  const double *bond_lib;              /* [][4] Dmz, amz, r0, fk (qbondlib)           */
  const int32_t *angle_ijk, *angle_cod;
  const double *angle_lib;             /* [][4] fk, ang0, ureyfk, ureyr0 (qanglib)    */
  const int32_t *torsion_ijkl, *torsion_cod;
  const double *torsion_lib;           /* [][3] fk, rmult, deltor (qtorlib)           */
  const int32_t *improper_ijkl, *improper_cod;
  const double *improper_lib;          /* [][2] fk, imp0 (qimplib)                    */
  /* Morse couplings [n][3]: term, bond, flag (iang_coupl / itor_coupl /
   * iimp_coupl(1:3, :), src/qatom.f90:1688-1800); flag 1 scales the term by
   * 1 - EMorseD(bond) instead of EMorseD(bond), any other value as 0 */
  int32_t nangle_coupl, ntorsion_coupl, nimproper_coupl;
  const int32_t *angle_coupl, *torsion_coupl, *improper_coupl;
  int32_t urey_bradley; /* 1: qurey_bradley after qangle (CHARMM, src/potene.f90:187) */
} qnb_qbonded;

/* Capacity of one Q bonded set (terms of all four kinds, and Q bonds). */
#define QNB_QB_MAX_TERMS 4096
#define QNB_QB_MAX_BONDS 1024

/* Loads (or with qb = NULL / no terms, removes) the Q bonded terms of a
 * handle and zeroes their Morse coupling state, as qatom_load_fep does
 * (EMorseD = dMorse = 0, src/qatom.f90:1000-1004). */
int32_t qnb_set_qbonded(qnb_handle *h, const qnb_qbonded *qb);

/* One pass of pot_energy's per-state Q bonded loop (src/potene.f90:183-190):
 * for each state qbond, qangle, [qurey_bradley], qtorsion, qimproper
 * (src/bondene.f90:387-685) at x, ADDING the lambda-weighted gradient into d
 * (natom x 3, on the host) and writing eq[s*4 + {0,1,2,3}] = EQ(s)%q%bond,
 * %angle (Urey-Bradley included), %torsion, %improper.  A coupled term reads
 * the Morse state (EMorseD, dMorse) its bond left at its last evaluation,
 * in this call or an earlier one, exactly as the reference's module arrays
 * carry it. */
int32_t qnb_eval_qbonded(qnb_handle *h, const double *x, double *d, double *eq);

/* ------------------------------------------------------------------------
 * Device-resident MD over a batch of instances (lambda windows of one system,
 * or independent systems of uneven size).  The step follows md_run
 * (src/md.f90:1657-1871) for leap-frog + Berendsen: nonbonded engine (list
 * update every nbcycle steps), solute bonds and angles, newvel_ber
 * (src/md.f90:1451-1466), SHAKE of the rigid waters (src/bondene.f90:
 * 1069-1151), v = (x - xx)/dt, temperature() (src/md.f90:1379-1447) which
 * gives the next step's scaling factor and Tfree.  All instances advance in
 * the same kernel launches; a list-update cycle is captured once as a HIP
 * graph and replayed.  Energies are sampled on the device every ene_every
 * steps into a per-instance history. */
typedef struct qnb_md_params {
  double dt_fs;       /* time step in fs; internal dt = 0.020462*dt_fs (src/md.f90:207) */
  double T0;          /* bath temperature (K)                                   */
  double tau_T_fs;    /* Berendsen coupling time (fs)                           */
  int32_t nbcycle;    /* pair-list update interval (NBcycle)                    */
  int32_t ene_every;  /* energy sampling interval (iene_cycle)                  */
  int32_t max_frames; /* capacity of the energy history                         */
  int32_t nbonds, nangles;
  const int32_t *bond_ij;   /* [nbonds][2]                                        */
  const double *bond_r0, *bond_k;     /* E = k/2 (r - r0)^2 (src/bondene.f90:182-244) */
  const int32_t *angle_ijk; /* [nangles][3], j the vertex                          */
  const double *angle_t0, *angle_k;   /* radians; E = k/2 (t - t0)^2            */
  const double *winv;       /* [natom] inverse masses                            */
  double water_roh, water_rhh; /* rigid 3-site water geometry (SHAKE constraints) */
  int32_t separate_scaling;    /* 1: solute and solvent are separate Berendsen groups
                                  (Qdyn6's default, src/md.f90:303, src/simprep.f90:3908-3923);
                                  0: one group for all atoms                          */
} qnb_md_params;

/* per-frame record of the energy history (float64 each) */
#define QNB_FRAME_LEN 104
enum qnb_frame_field {
  QNB_F_PP_EL = 0, QNB_F_PP_VDW, QNB_F_PW_EL, QNB_F_PW_VDW, QNB_F_WW_EL, QNB_F_WW_VDW, QNB_F_LRF,
  QNB_F_RAD, QNB_F_POL,
  QNB_F_Q0 = 9,          /* per state s: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw at 9 + 6s (8 states) */
  QNB_F_BOND = 57, QNB_F_ANGLE, QNB_F_EKIN, QNB_F_TEMP, QNB_F_TFREE, QNB_F_STEP,
  QNB_F_OFFD = 63,       /* EVB off-diagonal m: Hij at 63 + 2m, r_kl at 64 + 2m (m < QNB_MAX_OFFD) */
  QNB_F_QB0 = 71         /* per state s: Q bond, angle, torsion, improper at 71 + 4s (8 states) */
};
#define QNB_MAX_OFFD 4

typedef struct qnb_batch qnb_batch;

/* n instances; instance k runs system sys[k] with params par[k] (its own
 * lambda, SCAAS constants) and MD params md[k] from x0[k], v0[k]
 * (natom*3 each, velocities in A per internal time unit).  Instances whose
 * sys[k] pointers are equal share one device copy of the topology. */
int32_t qnb_batch_create(int32_t n, const qnb_system *const *sys, const qnb_params *const *par,
                         const qnb_md_params *const *md, const double *const *x0, const double *const *v0,
                         qnb_batch **out);
int32_t qnb_batch_destroy(qnb_batch *b);
/* Advances every instance by nsteps MD steps (device-resident).  With
 * use_graph, each list-update cycle of nbcycle steps is one HIP-graph launch;
 * the steps before the first cycle boundary and after the last one run as
 * ordinary launches.  nsteps = 0 with use_graph captures and instantiates the
 * cycle graph without advancing (so a timed run does not include it). */
int32_t qnb_batch_run(qnb_batch *b, int32_t nsteps, int32_t use_graph);
/* One integration step of every instance from a caller-supplied gradient:
 * d[k] (natom*3) replaces instance k's gradient, then leap-frog with each
 * temperature group's Berendsen factor (newvel_ber, src/md.f90:1451-1466),
 * SHAKE of the rigid waters and v = (x - x_old)/dt (src/md.f90:1787-1793,
 * src/bondene.f90:1069-1151) and temperature() for the next factor
 * (src/md.f90:1379-1447): the part of md_run's step after pot_energy, for a
 * caller that computes the forces itself.  The step counter advances. */
int32_t qnb_batch_integrate(qnb_batch *b, const double *const *d);
/* Thermostat state of instance k after the last step (temperature(),
 * src/md.f90:1379-1447): out[0..7] = Temp, Tfree, kinetic energy, tfree of
 * groups 1 and 2, Berendsen factor of groups 1 and 2 for the next step,
 * number of groups (1 or 2; unused group entries are 0). */
int32_t qnb_batch_thermostat(qnb_batch *b, int32_t k, double *out);
/* Waits for queued work; returns QNB_ENONFINITE / QNB_EOVERFLOW if an instance
 * went non-finite, a constraint failed or a list overflowed. */
int32_t qnb_batch_sync(qnb_batch *b);
int32_t qnb_batch_get_state(qnb_batch *b, int32_t k, double *x, double *v);
/* Copies instance k's energy history (nframes x QNB_FRAME_LEN float64). */
int32_t qnb_batch_energy_history(qnb_batch *b, int32_t k, double *out, int32_t cap, int32_t *nframes);
/* Pair interactions of instance k's current lists, counted as the reference's
 * explicit lists would hold them: out[0..6] = nbpp, nbpw, nbww, nbqp, nbqw
 * (atom pairs, one state), nbqq, nbqqp (summed over states). */
int32_t qnb_batch_pair_counts(qnb_batch *b, int32_t k, int64_t *out);

/* EVB off-diagonal elements sampled with the energies (offdiag,
 * src/nonbondene.f90:6238-6260): element m of instance k is
 * H = A[m] exp(-mu[m] r), r the distance between Q-atoms kl[2m] and
 * kl[2m+1] (0-based Q-atom numbers), and each sampled frame holds (H, r) at
 * QNB_F_OFFD + 2m, as Qdyn6 stores offd(:)%Hij, %rkl for the energy file.
 * n <= QNB_MAX_OFFD; n = 0 removes them. */
int32_t qnb_batch_set_offdiag(qnb_batch *b, int32_t k, int32_t n, const int32_t *kl, const double *A,
                              const double *mu);

/* Q bonded terms of instance k (qnb_qbonded above; NULL removes them): each
 * MD step adds their lambda-weighted forces (potene.f90:183-190) and every
 * sampled frame holds the per-state energies at QNB_F_QB0 + 4s.  Zeroes the
 * instance's Morse coupling state. */
int32_t qnb_batch_set_qbonded(qnb_batch *b, int32_t k, const qnb_qbonded *qb);

/* Relaxes the current structures before MD: nsteps steepest-descent moves
 * along the full gradient (nonbonded engine + solute bonds/angles), each atom
 * moving at most max_disp A per step, waters kept rigid by SHAKE.  Set-up aid
 * for synthetic starts (Qdyn6 relies on a prepared structure instead). */
int32_t qnb_batch_minimize(qnb_batch *b, int32_t nsteps, double max_disp);

/* Autotuner: picks the launch shape (charge groups = wavefronts per block,
 * 1/2/4/8) of the list/LRF sweep and of the force kernel by timing each
 * candidate `reps` times on the live batch; wpb_out[2] receives the choice
 * (lists, forces).  Call before qnb_batch_run; it drops a captured graph. */
int32_t qnb_batch_autotune(qnb_batch *b, int32_t reps, int32_t *wpb_out);

/* Times one kernel of the step on the batch's stream with HIP events: which =
 * 0 the list/LRF update (k_lists_lrf); 1 the receiver-side pair forces
 * (k_wforces for water receivers, k_sforces for solute receivers, in the form
 * used on steps that sample no energies) followed by the per-atom
 * LRF/restraint kernel; 2 k_wforces alone; 3 k_sforces alone; 4 the Q-atom
 * side of the Q-solute and Q-water pairs (k_qforces; QNB_EINVAL without
 * Q-atoms).  The kernels are launched `reps` times over all instances and
 * *ms_per_launch receives the average launch duration.  The positions,
 * velocities and recorded energies do not change; the gradient is left with
 * the timed launches' sums (k_qforces adds into it) and is recomputed by the
 * next step's force pass before anything reads it. */
int32_t qnb_batch_time_kernel(qnb_batch *b, int32_t which, int32_t reps, double *ms_per_launch);

/* ------------------------------------------------------------------------
 * Adjoint of the per-state Q-atom energies E_s = EQ(s)%qq + %qp + %qw
 * (src/potene.f90:224-229) over stored frames, for gradients of energy gaps
 * and Zwanzig/EVB free energies (src/qfep.f90:402-827) with respect to the
 * Q-side parameters that EVB calibration fits.  No counterpart in the
 * reference (Qfep6 has no derivatives).
 *
 * Every Q table entry (qp_*, qw_*, qq_* of qnb_system) is a product of one
 * Q-atom's parameter in one state and a partner factor
 * (precompute_set_values_qp/qq/qw, src/simprep.f90:3420-3589).  The caller
 * passes, per entry, the derivative of the entry with respect to that Q-side
 * parameter: for the charge, d el/d q_iq (the partner's scaled charge, times
 * el14_scale for 1-4 pairs); for LJ, d A/d A_iq and d B/d B_iq (with the
 * geometric rule the partner's A and B factors of the pair's LJ code).  qq
 * entries carry both members' tangents.  Entries that do not interact take
 * zeros.  Layouts are those of the matching qnb_system tables. */
typedef struct qnb_qtangent {
  const double *qp_dq, *qp_dA, *qp_dB;   /* [j + nat_solute*(iq + nqat*s)]       */
  const uint8_t *qp_is14;                /* 1: the pair is 1-4 (separate slots)   */
  const double *qw_dq, *qw_dA, *qw_dB;   /* [iq + nqat*(k + solv_atom*s)]         */
  const double *qq_dq_i, *qq_dA_i, *qq_dB_i; /* [iq + nqat*(jq + nqat*s)], wrt iq */
  const double *qq_dq_j, *qq_dA_j, *qq_dB_j; /* same entries, wrt jq              */
  const uint8_t *qq_is14;
} qnb_qtangent;

/* per (frame, state, Q-atom) output record, float64 each */
#define QNB_ADJ_NV 7
enum qnb_adj_field {
  QNB_ADJ_DQ = 0,   /* dE_s / d q_iq,s   (scaled charge)                  */
  QNB_ADJ_DA = 1,   /* dE_s / d A_iq,s   normal pairs                     */
  QNB_ADJ_DB = 2,   /* dE_s / d B_iq,s   normal pairs                     */
  QNB_ADJ_DA14 = 3, /* dE_s / d A_iq,s   1-4 pairs                        */
  QNB_ADJ_DB14 = 4, /* dE_s / d B_iq,s   1-4 pairs                        */
  QNB_ADJ_EEL = 5,  /* iq's share of E_s: electrostatic                   */
  QNB_ADJ_EVDW = 6  /* iq's share of E_s: van der Waals                   */
};

/* Host-supplied frames: frames[nframes][natom*3]; out[nframes][nstates][nqat]
 * [QNB_ADJ_NV].  The Q lists are rebuilt per frame from Rcq and xpcent;
 * pair math runs in the handle's precision (QNB_PREC_MIXED: FP32 pairs, FP64
 * sums), every sum in FP64.  Summing QNB_ADJ_EEL + QNB_ADJ_EVDW over Q-atoms
 * gives E_s of the frame. */
int32_t qnb_adjoint(qnb_handle *h, const qnb_qtangent *tan, int32_t nframes, const double *frames, double *out);

/* qnb_adjoint for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and results as qnb_adjoint;
 * the tangent tables, a chunk of frames, the kernel's parameter record and
 * the chunk's output are staged in a fixed device workspace of
 * QNB_ADJOINT_STATIC_BYTES per GPU (held until the process ends) instead of
 * allocated memory, so no call allocates or frees.  Returns QNB_ENOMEM, with
 * the sizes in the message, when the tangent tables plus one frame and its
 * output do not fit.  Every call on one GPU shares that GPU's workspace:
 * calls that overlap on one GPU race. */
#define QNB_ADJOINT_STATIC_BYTES (32u << 20)
int32_t qnb_adjoint_static(qnb_handle *h, const qnb_qtangent *tan, int32_t nframes, const double *frames,
                           double *out);

/* Coordinate snapshots during qnb_batch_run: every `every` steps (at the
 * point where the step's energies are sampled) up to max_frames per
 * instance, kept in device memory.  every = 0 stops storing. */
int32_t qnb_batch_store_frames(qnb_batch *b, int32_t every, int32_t max_frames);
/* Copies instance k's stored frames (nframes x natom*3 float64). */
int32_t qnb_batch_frames(qnb_batch *b, int32_t k, double *out, int32_t cap, int32_t *nframes);
/* Adjoint over every instance's stored frames in one launch, in the batch's
 * precision.  tan[k]: tangents of instance k (pointers may repeat).  out
 * receives the instances' records one after another (instance k:
 * nframes_k x nstates_k x nqat_k x QNB_ADJ_NV); *n the number of float64
 * written (also when it exceeds cap, then nothing is copied). */
int32_t qnb_batch_adjoint(qnb_batch *b, const qnb_qtangent *const *tan, double *out, int64_t cap, int64_t *n);

/* ------------------------------------------------------------------------
 * Group contributions (calculate_exc, src/nonbondene.f90:6825-6920): energy
 * arrays that repeat the per-state Q-atom energies with the interactions
 * between the Q-atoms and a group of solute atoms removed, as Qdyn6 writes
 * them after the first array of an energy file for the input's
 * [group_contribution] lines (src/md.f90:1011-1070; a line of type "all"
 * gives three arrays: full, electro, vdw).  Array a: mask[a*nat_solute + i]
 * = 1 for the group's solute atoms, type[a] one of QNB_GC_*.  The pairs are
 * those exc_assign lists (src/simprep.f90:2868-2947) from the Q tables of
 * `sys` (the system the engine was created from), with no cut-off.  Results
 * per array a and state s: QNB_GC_NV float64 = qq el, qq vdW, qp el, qp vdW,
 * qx el, qx vdW, total. */
enum qnb_gc_type { QNB_GC_FULL = 1, QNB_GC_ELECTRO = 2, QNB_GC_VDW = 3 };
#define QNB_GC_NV 7
/* Sets (narr = 0: removes) the engine's arrays. */
int32_t qnb_set_groups(qnb_handle *h, const qnb_system *sys, int32_t narr, const uint8_t *mask, const int32_t *type);
/* calculate_exc at x (natom*3, host) from the per-state energies pot_energy
 * left: eq_in[s*11 + 0..10] = qq el, qq vdW, qp el, qp vdW, qw el, qw vdW, Q
 * bond, angle, torsion, improper, restraint of state s; out[(a*nstates + s)
 * * QNB_GC_NV + k].  Pair by pair in the reference's order in FP64. */
int32_t qnb_calculate_exc(qnb_handle *h, const double *x, const double *eq_in, double *out);
/* The same for batch instance k, computed on every sampled step from the
 * step's energy record; frames recorded before the arrays were set hold 0. */
int32_t qnb_batch_set_groups(qnb_batch *b, int32_t k, const qnb_system *sys, int32_t narr, const uint8_t *mask,
                             const int32_t *type);
/* Copies instance k's group energies: nframes x narr x nstates x QNB_GC_NV
 * float64 (nothing when *nframes exceeds cap frames). */
int32_t qnb_batch_group_energies(qnb_batch *b, int32_t k, double *out, int32_t cap, int32_t *nframes);

/* ------------------------------------------------------------------------
 * Free-energy analysis on the device: Qfep6's Part 1 (Zwanzig forward and
 * reverse averages and the dG chain, src/qfep.f90:587-648), Part 2 (EVB
 * ground-state reaction free energy and diabatic curves per window and
 * energy-gap bin, :700-795, two states) and Part 3 (bin averages over the
 * windows, :797-827), every sum over the frames in the reference's order.
 * Replaces running Qfep6 on the energy files of a batch. */
typedef struct qnb_fep_params {
  double kT;                       /* R T in kcal/mol (Qfep6's rt)              */
  int32_t nskip;                   /* frames skipped at the start of a window   */
  int32_t nbins;                   /* energy-gap bins; 0: Part 1 only           */
  int32_t nmin;                    /* fewest points of a reported bin           */
  int32_t reserved;
  double alpha[QNB_MAX_STATES];    /* alpha shifts (alpha[0] is taken as 0)     */
  double coeff[QNB_MAX_STATES];    /* energy-gap coefficients                   */
  /* H12 = A exp(-mu (r - r0)) exp(-eta (r - r0)^2) (qfep.f90:476-480), r the
   * frame's distance of the first off-diagonal, 0 where none is recorded;
   * c1 is computed only where off-diagonals are recorded, as in Qfep6 */
  double A, mu, eta, r0;
} qnb_fep_params;

/* Results, caller-owned host arrays (any may be NULL); n windows, B bins,
 * S states.  Unreported entries are NaN. */
typedef struct qnb_fep_out {
  double *dgf;        /* [n] window i -> i+1 (i < n-1)                        */
  double *dgr;        /* [n] window i -> i-1 (i > 0)                          */
  double *dG;         /* [n] dG_0 = 0                                         */
  double *gap_range;  /* [2] smallest and largest gap after nskip             */
  double *dGg;        /* [n][B] ground-state free energy per window and bin   */
  double *dGv;        /* [n][B][S] diabatic free energies                     */
  double *c1, *c2;    /* [n][B] mean c1^2, c2^2                               */
  double *r;          /* [n][B] mean first off-diagonal distance              */
  int32_t *npts;      /* [n][B] points in the bin                             */
  double *bin_dGg, *bin_c1, *bin_c2, *bin_r; /* [B] Part 3 averages          */
  int32_t *ptsum;     /* [B] points over the reported windows                 */
} qnb_fep_out;

/* Over a batch's recorded energy frames, one window per instance in order
 * (lambdas from the instances; every instance must have the same number of
 * states and more than nskip frames).  The state energies are formed from the
 * frames as pot_energy forms EQ(s)%total (src/potene.f90:224-231). */
int32_t qnb_batch_free_energy(qnb_batch *b, const qnb_fep_params *p, qnb_fep_out *out);
/* Over state energies the caller holds (e.g. read from Qdyn6 energy files):
 * window w has nframes[w] frames of nstates EQ(s)%total values at V[w] (host,
 * frame-major), optionally rkl[w][nframes[w]] (the first off-diagonal's
 * distance; rkl or rkl[w] NULL: none) and lambdas[w*nstates .. +nstates). */
int32_t qnb_free_energy(int32_t nwin, int32_t nstates, const int32_t *nframes, const double *const *V,
                        const double *const *rkl, const double *lambdas, const qnb_fep_params *p, qnb_fep_out *out);

#ifdef __cplusplus
}
#endif
#endif /* QNB_H */
