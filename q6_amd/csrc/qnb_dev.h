// qnb_dev.h -- device-side data layout and pair arithmetic of the MI355X
// nonbonded engine (gfx950, wave64).
//
// Layout in HBM (one record per system instance; topology arrays may be shared
// between instances of a batch, per-instance state is private):
//   x, d        natom x {x,y,z} float64 (reference qr_vec layout)
//   lrf         ncgp x 43 float64 (centre, phi0, phi1, phi2(3x3), phi3(9x3),
//               the reference's LRF_TYPE order)
//   nbr/nnbr    full neighbour lists in ELL form: row g holds every charge group
//               within the cut-off of group g (both directions of a pair),
//               nnbr[g] entries, row stride `cap`
//   epart       per-wave energy partials [slot][QNB_NE] reduced in fixed order
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qnb.h"

#define QNB_WAVE 64
#define QNB_NLRF 43
#define QNB_MAXST 8
#define QNB_MAXSH 8

// energy slots
enum {
  E_PP_EL = 0, E_PP_VDW, E_PW_EL, E_PW_VDW, E_WW_EL, E_WW_VDW, E_LRF, E_RAD, E_POL,
  E_Q0  // then per state s: qq_el, qq_vdw, qp_el, qp_vdw, qw_el, qw_vdw at E_Q0 + 6*s
};
#define QNB_NE (E_Q0 + 6 * QNB_MAXST)

struct DevSys {
  // sizes
  int natom, nat_solute, nwat, sa, ncgp, ncgp_solute, nqat, nstates, ntypes, ivdw_rule, spc;
  // 1: every solute-water LJ parameter of solvent atoms 1.. is zero (their
  // pw pairs reduce exactly to Coulomb)
  int pw_coul_h;
  int any_atom;  // iuse_switch_atom = 0: group cut-off tests take any atom pair
  int use_lrf, prec, quirk;
  double rcpp2, rcpw2, rcww2, rcq2, rclrf2, rexcl2;
  double lambda[QNB_MAXST];
  double el14;
  double xp[3], xw[3];
  // SCAAS
  double rwat, fk_wsphere, Dwmz, awmz, Boltz, fkwpol, itdis;
  int wpol, nshell;
  double sh_rout[QNB_MAXSH], sh_dr[QNB_MAXSH], sh_cstb[QNB_MAXSH];

  // topology (may be shared across a batch)
  const double* crg;
  const int* iac;
  const int* iqatom;
  const uint8_t* excl;
  const uint8_t* heavy;
  const int* gfirst;   // into cgpatom
  const int* glast;    // into cgpatom (inclusive)
  const int* gswitch;  // atom
  const int* cgpatom;
  const int* atom_group;
  // per group-ordered atom k (cgpatom order; groups are consecutive runs):
  // the atom's group modulo 64, its bit in the far-group ballot word of the
  // list sweep's 64-group chunk
  const int* gslot;
  int ncgpat;
  const double* avdw;  // [t*3+k]
  const double* bvdw;
  const int* ljcod;    // [ti + tj*ntypes], 1-based code
  const uint32_t* exmask;  // [nat_solute] bit k-1: pair (i, i+k) excluded; bit 31: i has long partners
  const uint32_t* m14mask; // [nat_solute] bit k-1: pair (i, i+k) is 1-4
  const int* lx_off;       // [nat_solute+1] long exclusion/1-4 partners of the lower atom
  const int* lx_j;
  const uint8_t* lx_kind;  // 0 excluded, 1 1-4
  // solute-receiver fast path (k_sforces): a partner group b whose atom
  // indices lie more than 25 from every atom of the receiver group a, with
  // no long-range exclusion or 1-4 partner in either group, holds no
  // excluded or 1-4 pair with a; with the geometric rule and LJ code 1 for
  // every type pair, A, B and qq of such a pair are products of per-atom
  // factors
  const int* gspan;       // [ncgp_solute][2]: smallest and largest atom index of the group
  const uint8_t* glong;   // [ncgp_solute]: an atom of the group has long-range exclusion / 1-4 partners
  int ppfast;             // geometric rule and ljcod = 1 throughout
  const double* ppd;      // [nat_solute][4]: avdw(t, 1), bvdw(t, 1), charge, 0 (FP64 path)
  const float* ppf;       // [nat_solute][4]: the same in FP32 (mixed path)
  const double* pwA;       // [i + j*nat_solute]
  const double* pwB;
  const double* pwEl;
  const float* pw_f;  // mixed path: [i][PW_REC] A, B, el of atom i with solvent sites 0..2 (FP32)
  const double* wwA;       // [la + ka*sa]
  const double* wwB;
  const double* wwEl;
  // pairs inside each solvent molecule (nonbond_solvent_internal): offsets
  // [n_sint][2] and A, B, charge product [n_sint][3]
  int n_sint;
  const int* sint_ij;
  const double* sint_par;
  const int* iqseq;
  const uint8_t* qconn3;
  const double* qpA;  // [i + nsol*(iq + nq*s)]
  const double* qpB;
  const double* qpEl;
  const double* qpSc;
  const float* qpA_f;  // FP32 copies of qpA, qpB, qpEl, qpSc for the mixed path
  const float* qpB_f;
  const float* qpEl_f;
  const float* qpSc_f;
  const double* qwA;  // [iq + nq*(j + sa*s)]
  const double* qwB;
  const double* qwEl;
  const double* qwSc;
  // static Q-Q and Q-bonded-neighbour lists, all states concatenated
  const int* qq_iq;
  const int* qq_jq;
  const int* qq_st;
  const uint8_t* qq_soft;
  const double* qq_par;  // [p*4]: A, B, el, score
  int nqq;
  const int* qqp_iq;
  const int* qqp_j;
  const int* qqp_st;
  const double* qqp_par;
  int nqqp;
  // atom-centred view of the two lists (Q-atoms, then the distinct partners
  // of the Q-neighbour pairs): items {first atom, second atom, pair,
  // (state << 2) | (kind << 1) | side}
  const int* qe_atom;
  const int* qe_off;
  const int* qe_item;  // [4 * items]
  int nqe;

  // per-instance state
  double* x;
  // LRF sources in cgpatom order, packed before each list update: x, y, z
  // and the charge (0 for Q-atoms, which take no part in the expansion);
  // FP32 for the mixed path, FP64 for the plain one (the other is null)
  float* xqg;
  double* xqgd;
  // [ncgp][4] per group, packed with them: its reference atom's position
  // (switch atom of a solute group, oxygen of a water) and 1 if that atom is
  // excluded, else 0
  double* gref;
  double* d;
  double* lrf;
  int* nbr;
  int* nnbr;
  // reference test-point reuse of the LRF-off switch-atom solute-water list
  // (quirk bit 0 with use_lrf = 0): per solute group, the first water within
  // the cut-off of the switch atom (nwat when none); later waters are tested
  // against atom glast + 1
  int* pw_first;
  int cap;
  int* need;        // [1] max neighbours seen (overflow detection)
  uint8_t* qw_in;   // [nwat] water on the Q-water list
  uint8_t* qp_in;   // [ncgp_solute] group on the Q-solute list
  int* qw_list;     // compacted waters on the Q list
  int* qp_list;     // compacted solute atoms on the Q list
  int* nq_lists;    // [2]: n_qw, n_qp
  double* epart;    // [nslots][QNB_NE]
  int nslots;
  double* ered;     // [QNB_NE]
  double* eatom;    // [natom][2]: per-atom LRF and radial-restraint energies
  // polarisation restraint state
  double* sh_state;  // [3*QNB_MAXSH]: theta_corr, avtheta, avn_insh
  double* w_theta;   // [nwat]
  int* w_shell;      // [nwat]
  int* w_rank;       // [nwat] rank of the water in its shell
  int* sh_list;      // [QNB_MAXSH][nwat] members of each shell
  double* sh_lth;    // [QNB_MAXSH][nwat] their theta
  int* sh_n;         // [QNB_MAXSH] members per shell
  double* wp_part;   // [ceil(nwat / WP_THREADS)][1 + QNB_MAXSH] per-block sums
  // energy sampling: energies are accumulated when ene_every <= 0 or when
  // the instance's step counter is a multiple of ene_every
  const int* step_ptr;
  int ene_every;
};

// Q-atom bonded terms of one instance (qbond, qangle, qurey_bradley,
// qtorsion, qimproper, src/bondene.f90:387-685), resolved per state on the
// host.  Terms are numbered bonds, angles, torsions, impropers; each writes
// its gradient contributions into fixed slots (zeros where absent), and every
// atom the terms touch sums its slots in a fixed order.
struct QbDev {
  int nb, na, nt, ni, ns, urey;
  int nslot, nat;
  const int* atoms;     // term atoms: 2 per bond, 3 per angle, 4 per torsion / improper
  const double* par;    // [nterm][ns][4] the state's library row
  const uint8_t* pres;  // [nterm][ns] 1 = term present in the state
  const int* cpl;       // [na + nt + ni][4]: bond, flag, Urey-Bradley bond, flag (bond -1 = uncoupled)
  const int* slot0;     // [nterm] first slot of the term
  const int* coff;      // [nat + 1] slots of distinct atom a: cslot[coff[a] .. coff[a+1])
  const int* cslot;
  const int* catom;     // [nat] atom number
  double* morse;        // [nb][4] EMorseD, dMorse_i (x, y, z): kept between calls
  double* snap;         // [nb][ns][4] the same as evaluated in each state of this call
  double* eterm;        // [nterm][ns] energy of each term in each state
  double* slots;        // [ns][nslot][3] gradient shares of each state's terms
  double* e;            // [ns][4] bond, angle, torsion, improper of the last call
  double lambda[QNB_MAXST];
};

__device__ __forceinline__ bool want_energies(const DevSys& S) {
  return S.ene_every <= 0 || (*S.step_ptr % S.ene_every) == 0;
}

// ---------------------------------------------------------------- helpers

// load through the global address space: the DevSys arrays are reached
// through generic pointers, whose loads (flat) the compiler can neither
// separate from LDS traffic in its wait counts nor batch as freely
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// a Q-solute table in the pair arithmetic's type: on the mixed path its FP32
// copy (the same values the FP64 table gives after rounding to float)
template <typename R>
__device__ __forceinline__ const R* qp_tab(const double* d, const float* f) {
  if constexpr (sizeof(R) == 4) return f;
  else return d;
}

template <typename T>
using gptr_t = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr_t<T> gp(const T* p) {
  return (gptr_t<T>)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gpm(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void stg(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, QNB_WAVE);
  return v;
}

// Sum over the wavefront without LDS instructions, for call sites where every
// lane is active: DPP moves within each row of 16 lanes (quad_perm xor 1 and
// xor 2, row_half_mirror, row_mirror), then the four row sums are read with
// v_readlane and added in a fixed order, so the result (the same in every
// lane) is deterministic.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_all(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

// Sums N per-lane values over the wave at once ("transposed" reduction): a
// step pairs every lane with a partner and, while each lane still holds
// several values, the lane keeps half of them (the half its lane bit names)
// and adds the partner's copy of that half, so the number of values per
// lane halves each step.  Pairings: row_mirror (lane bit 3), row_half_mirror
// (bit 2), quad_perm xor 2 (bit 1), xor 1 (bit 0), then v_permlane16_swap
// (bit 4: rows 0<->1, 2<->3) and v_permlane32_swap (bit 5: halves); the
// mirrors come first because they also flip the lower bits, which must not
// yet decide anything.  Once a lane holds one value the remaining steps are
// plain butterflies.  Returns, in every lane, the full sum of value
// wave_sum_tr_index<N>(lane).  About 7 VALU per value (13% of N separate
// wave_sum_all reductions for N = 30).
__host__ __device__ constexpr int pow2_ceil(int n) { return n <= 1 ? 1 : 2 * pow2_ceil((n + 1) / 2); }
__host__ __device__ constexpr int log2_exact(int p) { return p <= 1 ? 0 : 1 + log2_exact(p / 2); }

template <int CTRL, int C>
__device__ __forceinline__ void tr_dpp_step(double* v, bool upper) {
  if constexpr (C >= 2) {
#pragma unroll
    for (int j = 0; j < C / 2; j++) {
      const double x = v[j], y = v[j + C / 2];
      const double send = upper ? x : y;  // the half the partner keeps
      const double keep = upper ? y : x;
      v[j] = keep + dpp_move<CTRL>(send);
    }
  } else {
    v[0] += dpp_move<CTRL>(v[0]);
  }
}

__device__ __forceinline__ void swap16(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                   false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap32(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                   false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}

// rows (lane bit 4) or halves (bit 5): after the swap, x + y holds in every
// lane the pair's sum of the half that lane keeps
template <int C, bool HALVES>
__device__ __forceinline__ void tr_swap_step(double* v) {
  if constexpr (C >= 2) {
#pragma unroll
    for (int j = 0; j < C / 2; j++) {
      double x = v[j], y = v[j + C / 2];
      if constexpr (HALVES) swap32(x, y);
      else swap16(x, y);
      v[j] = x + y;
    }
  } else {
    double x = v[0], y = v[0];
    if constexpr (HALVES) swap32(x, y);
    else swap16(x, y);
    v[0] = x + y;
  }
}

template <int CTRL>
__device__ __forceinline__ float dpp_move_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// per-lane partials of type T (float or double); the first step exchanges
// them in T and adds in FP64 ((double)a + (double)b, exact conversions), so
// float partials are never summed in float and the wide FP64 copy of the N
// values is never live at once
template <typename T, int N>
__device__ __forceinline__ double wave_sum_tr(const T (&in)[N], int lane) {
  constexpr int P = pow2_ceil(N);
  static_assert(P <= 64, "at most 64 values");
  static_assert(QNB_NE <= 64, "one energy record per wave store");
  double v[P];
  const bool up = lane & 8;
  if constexpr (P >= 2) {
#pragma unroll
    for (int j = 0; j < P / 2; j++) {
      const T x = in[j], y = j + P / 2 < N ? in[j + P / 2] : T(0);
      const T send = up ? x : y;
      const T keep = up ? y : x;
      T got;
      if constexpr (sizeof(T) == 4) got = dpp_move_f<0x140>(send);
      else got = dpp_move<0x140>(send);
      v[j] = (double)keep + (double)got;
    }
  } else {
    T got;
    if constexpr (sizeof(T) == 4) got = dpp_move_f<0x140>(in[0]);
    else got = dpp_move<0x140>(in[0]);
    v[0] = (double)in[0] + (double)got;
  }
  tr_dpp_step<0x141, (P >= 2 ? P / 2 : 1)>(v, lane & 4);
  tr_dpp_step<0x4E, (P >= 4 ? P / 4 : 1)>(v, lane & 2);
  tr_dpp_step<0xB1, (P >= 8 ? P / 8 : 1)>(v, lane & 1);
  tr_swap_step<(P >= 16 ? P / 16 : 1), false>(v);
  tr_swap_step<(P >= 32 ? P / 32 : 1), true>(v);
  return v[0];
}

// the value whose sum lane `lane` holds after wave_sum_tr<N>, and whether
// the lane is that value's one writer (its bits that decided nothing are 0)
template <int N>
__device__ __forceinline__ int wave_sum_tr_index(int lane) {
  constexpr int L = log2_exact(pow2_ceil(N));
  const int bits[6] = {3, 2, 1, 0, 4, 5};
  int j = 0;
#pragma unroll
  for (int s = 0; s < L; s++) j = (j << 1) | ((lane >> bits[s]) & 1);
  return j;
}
template <int N>
__device__ __forceinline__ bool wave_sum_tr_writer(int lane) {
  constexpr int L = log2_exact(pow2_ceil(N));
  constexpr int bits[6] = {3, 2, 1, 0, 4, 5};
  int decided = 0;
#pragma unroll
  for (int s = 0; s < L; s++) decided |= 1 << bits[s];
  return (lane & ~decided) == 0 && wave_sum_tr_index<N>(lane) < N;
}

// the lane that holds value j after wave_sum_tr<N> and writes it
template <int N>
__host__ __device__ constexpr int wave_sum_tr_lane(int j) {
  constexpr int L = log2_exact(pow2_ceil(N));
  constexpr int bits[6] = {3, 2, 1, 0, 4, 5};
  int lane = 0;
  for (int s = 0; s < L; s++) lane |= ((j >> (L - 1 - s)) & 1) << bits[s];
  return lane;
}

// A wave's energy record ep[0 .. QNB_NE) in one coalesced store from lanes
// 0 .. QNB_NE-1: the N per-lane partials are summed with one transposed
// reduction, each sum is read from its lane, and entry slot[j] takes sum j,
// every other entry 0.
template <int N>
__device__ __forceinline__ void store_energy_record(double* ep, const double (&part)[N], const int (&slot)[N],
                                                    int lane) {
  const double sum = wave_sum_tr(part, lane);
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < N; j++) {
    const double vj = lane_value(sum, wave_sum_tr_lane<N>(j));
    if (lane == slot[j]) v = vj;
  }
  if (lane < QNB_NE) *(__attribute__((address_space(1))) double*)(ep + lane) = v;
}

// group contributions (calculate_exc): one pair of an energy array's list,
// kind 0 Q-Q (nbe_qq; soft: the exponential soft pair), 1 Q-solute (nbe_qx);
// i the Q-atom's index, j the partner's; the pair's A, B, el and soft-core
// score as the Q tables hold them
struct ExcEnt {
  int kind, i, j, soft;
  double A, B, el, sc;
};
// one system's arrays: off[a*ns + s] .. off[a*ns + s + 1] the pairs of array
// a in state s, type[a] (QNB_GC_*); hist: the batch's per-frame results
// [frame][narr][ns][QNB_GC_NV]
struct ExcDev {
  int narr, ns, max_frames;
  const int* off;
  const ExcEnt* ent;
  const int* type;
  double* hist;
};

// the walk of one (array, state): e[11] the state's energies (qq el, vdW, qp
// el, vdW, qw el, vdW, Q bond, angle, torsion, improper, restraint), o[7]
__device__ __forceinline__ void exc_walk(const ExcDev& D, const double* __restrict__ x, int a, int st,
                                         const double* e, double* o) {
#pragma clang fp contract(off)
  double qqel = e[0], qqvdw = e[1], qpel = e[2], qpvdw = e[3];
  const int ty = D.type[a];
  const int p0 = D.off[a * D.ns + st], p1 = D.off[a * D.ns + st + 1];
  for (int p = p0; p < p1; p++) {
    const ExcEnt& t = D.ent[p];
    const double ux = x[3 * t.j] - x[3 * t.i], uy = x[3 * t.j + 1] - x[3 * t.i + 1], uz = x[3 * t.j + 2] - x[3 * t.i + 2];
    const double r2 = 1.0 / (ux * ux + uy * uy + uz * uz), r = sqrt(r2), r6d = r2 * r2 * r2;
    const double r6_hc = 1.0 / r6d;
    double r6 = r6_hc + t.sc;
    r6 = 1.0 / r6;
    const double r12 = r6 * r6;
    const double Vel = t.el * r;
    double Va, Vb;
    if (t.soft) {  // nbe_qq's exponential soft pair (nonbonded.f90:138-141)
      Vb = 0.0;
      Va = t.A * exp(-t.B / r);
    } else {
      Va = t.A * r12;
      Vb = t.B * r6;
    }
    double* el = t.kind == 0 ? &qqel : &qpel;
    double* vdw = t.kind == 0 ? &qqvdw : &qpvdw;
    if (ty == QNB_GC_FULL) {
      *el = *el - Vel;
      *vdw = *vdw - (Va - Vb);
    } else if (ty == QNB_GC_ELECTRO) {
      *el = *el - Vel;
    } else {
      *vdw = *vdw - (Va - Vb);
    }
  }
  o[0] = qqel;
  o[1] = qqvdw;
  o[2] = qpel;
  o[3] = qpvdw;
  o[4] = qqel + qpel + e[4];
  o[5] = qqvdw + qpvdw + e[5];
  o[6] = e[6] + e[7] + e[8] + e[9] + o[4] + o[5] + e[10];
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// the LRF-off switch-atom solute-water list takes nbpwlist's test-point reuse
__host__ __device__ __forceinline__ bool pw_quirk(const DevSys& S) {
  return (S.quirk & 1) && !S.use_lrf && !S.any_atom && S.nwat > 0 && S.ncgp_solute > 0;
}

// q_dist4 (math.f90:254-262): |b - a|^2, ((dx^2 + dy^2) + dz^2), no contraction
__device__ __forceinline__ double dist4(const double* __restrict__ x, int a, int b) {
#pragma clang fp contract(off)
  double dx = ldg(x + 3 * b) - ldg(x + 3 * a);
  double dy = ldg(x + 3 * b + 1) - ldg(x + 3 * a + 1);
  double dz = ldg(x + 3 * b + 2) - ldg(x + 3 * a + 2);
  double r = dx * dx;
  r = r + dy * dy;
  r = r + dz * dz;
  return r;
}

__device__ __forceinline__ double dist4p(const double* __restrict__ x, int a, const double* p) {
#pragma clang fp contract(off)
  double dx = p[0] - x[3 * a];
  double dy = p[1] - x[3 * a + 1];
  double dz = p[2] - x[3 * a + 2];
  double r = dx * dx;
  r = r + dy * dy;
  r = r + dz * dz;
  return r;
}

// Adds an atom's per-atom terms to its gradient.  A solute atom in no charge
// group (an excluded atom of a topology cut to a sphere) is reached by no pair
// kernel, so this pass is the first writer of its gradient and stores it.
__device__ __forceinline__ void store_atom_gradient(const DevSys& S, int i, const double* f) {
  const auto d = gpm(S.d);
  if (i < S.nat_solute && ldg(S.atom_group + i) < 0) {
    d[3 * i] = f[0];
    d[3 * i + 1] = f[1];
    d[3 * i + 2] = f[2];
  } else {
    d[3 * i] += f[0];
    d[3 * i + 1] += f[1];
    d[3 * i + 2] += f[2];
  }
}

// LRF Taylor application for one atom (lrf_taylor, nonbondene.f90:507-571)
__device__ __forceinline__ void lrf_taylor_atom(const DevSys& S, int i, double* f, double& e) {
  const auto L = gp(S.lrf + (size_t)QNB_NLRF * ldg(S.atom_group + i));
  const auto x = gp(S.x);
  const double drx = L[0] - x[3 * i], dry = L[1] - x[3 * i + 1], drz = L[2] - x[3 * i + 2];
  const auto p1 = L + 4;
  const auto p2 = L + 7;
  const auto p3 = L + 16;
  double t0 = drx * p2[0] + dry * p2[1] + drz * p2[2];
  double t1 = drx * p2[3] + dry * p2[4] + drz * p2[5];
  double t2 = drx * p2[6] + dry * p2[7] + drz * p2[8];
  double V = L[3] + (drx * p1[0] + dry * p1[1] + drz * p1[2]) + 0.5 * (drx * t0 + dry * t1 + drz * t2);
  const double q = ldg(S.crg + i);
  e += 0.5 * q * V;
  double df[3];
  for (int k = 0; k < 3; k++) {
    const auto P = p3 + 9 * k;
    double u0 = drx * P[0] + dry * P[1] + drz * P[2];
    double u1 = drx * P[3] + dry * P[4] + drz * P[5];
    double u2 = drx * P[6] + dry * P[7] + drz * P[8];
    df[k] = p1[k] + (drx * p2[3 * k] + dry * p2[3 * k + 1] + drz * p2[3 * k + 2]) +
            0.5 * (drx * u0 + dry * u1 + drz * u2);
  }
  f[0] -= df[0] * q;
  f[1] -= df[1] * q;
  f[2] -= df[2] * q;
}

// SCAAS radial restraint on one water oxygen (restrain_solvent,
// nonbondene.f90:6466-6544, 3-site branch)
__device__ __forceinline__ void radial_restraint(const DevSys& S, int src, double shift, double* f, double& e) {
  const auto x = gp(S.x);
  double vx = x[3 * src] - S.xw[0], vy = x[3 * src + 1] - S.xw[1], vz = x[3 * src + 2] - S.xw[2];
  double b = sqrt(vx * vx + vy * vy + vz * vz);
  double db = b - (S.rwat - shift);
  double erst, dv;
  if (db > 0.0) {
    erst = 0.5 * S.fk_wsphere * (db * db) - S.Dwmz;
    dv = S.fk_wsphere * db / b;
  } else if (b > 0.0) {
    double fexp = exp(S.awmz * db);
    erst = S.Dwmz * (fexp * fexp - 2.0 * fexp);
    dv = -2.0 * S.Dwmz * S.awmz * (fexp - fexp * fexp) / b;
  } else {
    erst = 0.0;
    dv = 0.0;
  }
  f[0] += vx * dv;
  f[1] += vy * dv;
  f[2] += vz * dv;
  e += erst;
}

// The atom whose distance from xwcent drives atom i's SCAAS radial term
// (restrain_solvent, nonbondene.f90:6466-6544), or -1: a 3-site water's first
// atom acts on itself; in other solvents every heavy atom (by the first
// molecule's pattern, is_heavy_solv) of a non-excluded molecule takes the
// term at the distance of the molecule's first atom, as the reference
// measures it.
__device__ __forceinline__ int radial_source(const DevSys& S, int i) {
  const int w = i - S.nat_solute;
  if (w < 0) return -1;
  const int isolv = w % S.sa, first = i - isolv;
  if (ldg(S.excl + first)) return -1;
  if (S.sa == 3) return isolv == 0 ? i : -1;
  return ldg(S.heavy + S.nat_solute + isolv) ? first : -1;
}

// Q-solute group test (nbqplist, nonbondene.f90:3647-3742): the group's
// switch atom within Rcq of xpcent; with any_atom (nbqplis2 :3408-3500) any
// atom of the group, scanned in group order
__device__ __forceinline__ bool qp_group_in(const DevSys& S, const double* __restrict__ x, int g) {
  const int sw = S.gswitch[g];
  if (S.excl[sw]) return false;
  if (!S.any_atom) return dist4p(x, sw, S.xp) <= S.rcq2;
  for (int k = S.gfirst[g]; k <= S.glast[g]; k++)
    if (dist4p(x, S.cgpatom[k], S.xp) <= S.rcq2) return true;
  return false;
}

// Solute-solute pair parameters computed from atom data exactly as
// pp_int_comp / precompute_set_values_pp (simprep.f90:2952-3065, 3321-3350)
// would tabulate them.  Returns false when the pair does not interact.
__device__ __forceinline__ bool pp_param(const DevSys& S, int i, int j, double& A, double& B, double& el) {
  int lo = i < j ? i : j, hi = i < j ? j : i, dk = hi - lo;
  bool is14 = false;
  if (dk <= 25) {
    uint32_t bit = 1u << (dk - 1);
    if (ldg(S.exmask + lo) & bit) return false;
    is14 = (ldg(S.m14mask + lo) & bit) != 0;
  } else {
    bool ex = false;
    for (int m = ldg(S.lx_off + lo); m < ldg(S.lx_off + (lo + 1)); m++)
      if (ldg(S.lx_j + m) == hi) {
        if (ldg(S.lx_kind + m) == 0) ex = true;
        else is14 = true;
      }
    if (ex) return false;
  }
  int ti = ldg(S.iac + i), tj = ldg(S.iac + j);
  int vdw = is14 ? 3 : ldg(S.ljcod + (ti + tj * S.ntypes));
  int k = vdw - 1;
  if (S.ivdw_rule == 1) {
    A = ldg(S.avdw + (ti * 3 + k)) * ldg(S.avdw + (tj * 3 + k));
    B = ldg(S.bvdw + (ti * 3 + k)) * ldg(S.bvdw + (tj * 3 + k));
  } else {
    double tA = ldg(S.avdw + (ti * 3 + k)) + ldg(S.avdw + (tj * 3 + k));
    tA = tA * tA;
    tA = tA * tA * tA;
    double tB = ldg(S.bvdw + (ti * 3 + k)) * ldg(S.bvdw + (tj * 3 + k));
    A = (tA * tA) * tB;
    B = 2.0 * tA * tB;
  }
  el = ldg(S.crg + i) * ldg(S.crg + j);
  bool set = fabs(A) > 1.0e-10 || fabs(B) > 1.0e-10 || fabs(el) > 1.0e-10;
  if (vdw == 3) el = el * S.el14;
  return set;
}

// 1/|u|^2 and 1/|u| from |u|^2: IEEE division and square root in FP64 (the
// plain path evaluates them as q_dist does); in FP32 the hardware reciprocal
// square root (v_rsq_f32, 1 ulp), the usual FP32 pair-kernel form
__device__ __forceinline__ void rinv(double d2, double& r2, double& r) {
  r2 = 1.0 / d2;
  r = sqrt(r2);
}
__device__ __forceinline__ void rinv(float d2, float& r2, float& r) {
  r = __builtin_amdgcn_rsqf(d2);
  r2 = r * r;
}
// 1/v: IEEE in FP64, hardware reciprocal (v_rcp_f32, 1 ulp) in FP32
__device__ __forceinline__ double recip(double v) { return 1.0 / v; }
__device__ __forceinline__ float recip(float v) { return __builtin_amdgcn_rcpf(v); }

// Pair arithmetic in precision R (double: plain path; float: mixed path).
// u = x_j - x_i arrives in FP64 (the difference of FP64 positions), is rounded
// to R, and the returned scalar dv is in R; callers accumulate u*dv in FP64.
template <typename R>
struct PairMath {
  // nbe (nonbonded.f90:45-75): LJ + Coulomb.  Adds V_el, V_a - V_b to the
  // accumulators (type T: FP64, or R for a short per-partner partial).
  // `want`: accumulate the energies (only the sampled steps need them)
  template <typename T>
  static __device__ __forceinline__ R nbe(R ux, R uy, R uz, R A, R B, R q, T& el, T& vdw, bool want = true) {
    R r2, r;
    rinv(ux * ux + uy * uy + uz * uz, r2, r);
    R r6 = r2 * r2 * r2;
    R r12 = r6 * r6;
    R Vel = q * r, Va = A * r12, Vb = B * r6;
    if (want) {
      el += (T)Vel;
      vdw += (T)Va - (T)Vb;
    }
    return r2 * (-Vel - R(12) * Va + R(6) * Vb);
  }
  // nbe_spc (nonbonded.f90:247-267): Coulomb only
  template <typename T>
  static __device__ __forceinline__ R nbe_spc(R ux, R uy, R uz, R q, T& el, bool want = true) {
    R r2, r;
    rinv(ux * ux + uy * uy + uz * uz, r2, r);
    R Vel = q * r;
    if (want) el += (T)Vel;
    return -r2 * Vel;
  }
  // nbe_qx (nonbonded.f90:200-222), soft-core score sc, state weight lam
  static __device__ __forceinline__ R nbe_qx(R r2, R r, R r6in, R A, R B, R q, R sc, R lam, double& el,
                                             double& vdw, bool want = true) {
    R r6_hc = recip(r6in);
    R r6 = recip(r6_hc + sc);
    R r12 = r6 * r6;
    R Vel = q * r, Va = A * r12, Vb = B * r6;
    if (want) {
      el += (double)Vel;
      vdw += (double)Va - (double)Vb;
    }
    return r2 * (-Vel - (R(12) * Va - R(6) * Vb) * r6 * r6_hc) * lam;
  }
};

// Receiver-side force kernel: at most FC_WPB wavefronts per block, FC_CAP
// staged partner atoms and FC_CH receiver atoms per sweep; per-wave LDS:
// partner offsets and receiver offsets (R x 3), partner index and tag,
// receiver index
// floats per solute atom in the FP32 solute-water table (3 sites x A, B, el,
// padded to 16-byte records)
#define PW_REC 12

// partners per wave of the Q-atom force kernel (the Q-atom's partners are
// split into chunks of about this many)
#ifndef QF_PART
#define QF_PART 512
#endif
// grid.x of the Q-atom force kernel: per XCD (8) and per Q-atom group of 8,
// the blocks of one Q-atom's chunks (W waves per block)
inline int qforces_blocks(int nq, int chunks, int W) { return 8 * ((chunks + W - 1) / W) * ((nq + 7) / 8); }

// threads of the Q-Q kernel's block (one block per system)
#define QQ_THREADS 1024
// threads per block of the polarisation-restraint kernels (one per water)
#define WP_THREADS 256
// threads of a polarisation-shell rank block (k_watpol_rank loops over any
// shell size with this stride; the launch uses the same value)
#ifndef WP_RANK_THREADS
#define WP_RANK_THREADS 1024
#endif

#define FC_WPB 8
#ifndef FC_CAP
#define FC_CAP 384
#endif
#ifndef FC_CH
#define FC_CH 10
#endif
// solute receivers: at most 168 VGPRs, so three waves fit a SIMD (the
// no-energy FP32 form fits; the energy-sampling and FP64 forms spill a few
// bytes to scratch)
#ifndef FC_OCC_ATTR
#define FC_OCC_ATTR __attribute__((amdgpu_waves_per_eu(3)))
#endif
template <typename R>
__host__ __device__ constexpr size_t fc_wave_bytes() {
  return ((sizeof(R) * (3 * (2 * FC_CAP + FC_CH) + 9 * FC_CH) + sizeof(double) * FC_CH +
           sizeof(int) * (2 * FC_CAP + 4 * FC_CH)) + 15) / 16 * 16;
}
// type tables of the solute-solute parameters held in LDS up to this many types
// Water-receiver force kernel (k_wforces): solute partner atoms are staged in
// rounds of at most CAP slots per wave (WF_CAP_SMALL when every solute group
// of the launch fits, else FC_CAP); partner waters need no staging.
#define WF_CAP_SMALL 128
// per-wave LDS of the solute-receiver kernel (k_sforces) with CAP partner
// slots: partner offsets, the sweep's receiver data, the partners' LJ and
// charge factors, then partner indices and tags and receiver records
template <typename R, int CAP>
__host__ __device__ constexpr size_t sf_wave_bytes() {
  return ((sizeof(R) * (3 * (2 * CAP + FC_CH) + 9 * FC_CH) + sizeof(double) * FC_CH +
           sizeof(int) * (2 * CAP + 4 * FC_CH)) + 15) / 16 * 16;
}
template <typename R, int CAP>
__host__ __device__ constexpr size_t wf_wave_bytes() {
  return (sizeof(R) * 3 * CAP + sizeof(int) * 2 * CAP + 15) / 16 * 16;
}

// Placement of a batch's blocks on the XCDs.  Blocks are dealt to the 8 XCDs
// round-robin in linear block order.  imap 0: grid (nb, ninst), block x of
// every instance on the same XCD (instances share the topology's rows in
// L2).  imap 1: a 1-D grid of 8 * ceil(nb * ninst / 8) blocks whose k-th
// block on XCD x takes work item x * per + k of the instance-major order, so
// each instance's blocks run on one or two XCDs and its coordinates stay in
// their L2.  Returns false for padding blocks.
__device__ __forceinline__ bool map_block(int imap, int nb, int ninst, int& inst, int& blk) {
  if (imap == 0) {
    inst = blockIdx.y;
    blk = blockIdx.x;
    return blk < nb;
  }
  const long total = (long)nb * ninst, per = (total + 7) / 8, L = blockIdx.x;
  const long g = (L % 8) * per + L / 8;
  if (g >= total) return false;
  inst = (int)(g / nb);
  blk = (int)(g % nb);
  return true;
}

// largest solute charge group of a system (gfirst/glast span)
inline int max_solute_group(const int* gfirst, const int* glast, int ncgp_solute) {
  int m = 0;
  for (int g = 0; g < ncgp_solute; g++) m = glast[g] - gfirst[g] + 1 > m ? glast[g] - gfirst[g] + 1 : m;
  return m;
}

#define FC_MAXT 64
// static LDS of the solute-receiver force kernel: the LJ factors and the
// ljcod table of up to FC_MAXT types
template <typename R>
__host__ __device__ constexpr size_t fc_static_bytes() {
  return 2 * 3 * FC_MAXT * sizeof(double) + FC_MAXT * FC_MAXT;
}

// Adjoint of the per-state Q-atom energies over a set of frames: tangent
// tables (derivative of each Q table entry with respect to the Q-side
// parameter of that entry) and the per-frame output.
#define QNB_ADJ_NV 7
enum { ADJ_DQ = 0, ADJ_DA, ADJ_DB, ADJ_DA14, ADJ_DB14, ADJ_EEL, ADJ_EVDW };

struct AdjDev {
  const double* frames;  // [nframes][natom*3]
  int nframes;
  const double* qp_dq;   // layout of qpA: [j + nsol*(iq + nq*s)]
  const double* qp_dA;
  const double* qp_dB;
  const uint8_t* qp_is14;
  const double* qw_dq;   // layout of qwA: [iq + nq*(k + sa*s)]
  const double* qw_dA;
  const double* qw_dB;
  const double* qq_dq_i;  // layout of the qq tables: [iq + nq*(jq + nq*s)]
  const double* qq_dA_i;
  const double* qq_dB_i;
  const double* qq_dq_j;
  const double* qq_dA_j;
  const double* qq_dB_j;
  const uint8_t* qq_is14;
  double* out;  // [nframes][nstates][nqat][QNB_ADJ_NV]
};
