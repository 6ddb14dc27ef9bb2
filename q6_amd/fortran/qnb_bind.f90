! qnb_bind.f90 -- ISO_C_BINDING interface of the MI355X nonbonded engine
! (include/qnb.h) for Fortran callers such as Qdyn6.
!
! The derived types mirror the C structs field for field (BIND(C)); the
! interfaces bind the C entry points by name.  A caller fills qnb_system with
! C_LOC of its (0-based where the header says so) arrays and keeps them alive
! for the duration of the call; qnb_create copies everything it needs.
module qnb_bind
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: QNB_OK = 0, QNB_PREC_FP64 = 0, QNB_PREC_MIXED = 1
  integer(c_int32_t), parameter :: QNB_MAX_STATES = 8, QNB_MAX_SHELLS = 8
  integer(c_int32_t), parameter :: QNB_T_WW = 1, QNB_T_QW = 2, QNB_T_PW = 4, QNB_T_PP = 8, QNB_T_QP = 16, &
                                   QNB_T_LRF = 32, QNB_T_QQ = 64, QNB_T_QQP = 128, QNB_T_RAD = 256, QNB_T_POL = 512, &
                                   QNB_T_SINT = 1024

  type, bind(c) :: qnb_system
    integer(c_int32_t) :: natom, nat_solute, nwat, solv_atom, ncgp, ncgp_solute, nqat, nstates
    integer(c_int32_t) :: ntypes, ivdw_rule, solvent_spc, nexlong, n14long, any_atom_cutoff
    type(c_ptr) :: crg, iac, iqatom, excl, heavy
    type(c_ptr) :: cgp_first, cgp_last, cgp_iswitch, cgpatom
    type(c_ptr) :: avdw, bvdw, ljcod
    real(c_double) :: el14_scale
    type(c_ptr) :: listex, list14, listexlong, list14long
    type(c_ptr) :: pw_A, pw_B, pw_el, ww_A, ww_B, ww_el
    type(c_ptr) :: iqseq, qconn3
    type(c_ptr) :: qp_A, qp_B, qp_el, qp_sc, qp_set
    type(c_ptr) :: qq_A, qq_B, qq_el, qq_sc, qq_set, qq_soft
    type(c_ptr) :: qw_A, qw_B, qw_el, qw_sc
    integer(c_int32_t) :: n_solv_int
    type(c_ptr) :: solv_int_ij, solv_int_par
  end type qnb_system

  type, bind(c) :: qnb_params
    real(c_double) :: Rcpp, Rcpw, Rcww, Rcq, RcLRF
    integer(c_int32_t) :: use_LRF, precision, ref_quirks
    real(c_double) :: xpcent(3), xwcent(3), rexcl_o
    real(c_double) :: lambda(QNB_MAX_STATES)
    real(c_double) :: rwat, fk_wsphere, Dwmz, awmz, Boltz
    integer(c_int32_t) :: wpol_restr, nwpolr_shell
    real(c_double) :: fkwpol, itdis
    real(c_double) :: wshell_rout(QNB_MAX_SHELLS), wshell_dr(QNB_MAX_SHELLS)
    real(c_double) :: wshell_cstb(QNB_MAX_SHELLS), wshell_theta_corr(QNB_MAX_SHELLS)
  end type qnb_params

  type, bind(c) :: qnb_energies
    real(c_double) :: pp_el, pp_vdw, pw_el, pw_vdw, ww_el, ww_vdw, lrf, solvent_radial, water_pol
  end type qnb_energies

  type, bind(c) :: qnb_qenergies
    real(c_double) :: qq_el, qq_vdw, qp_el, qp_vdw, qw_el, qw_vdw
  end type qnb_qenergies

  ! tangent tables of the Q tables for the adjoint (c_loc of real(8) /
  ! integer(1) arrays in the layouts of the matching qnb_system tables)
  type, bind(c) :: qnb_qtangent
    type(c_ptr) :: qp_dq, qp_dA, qp_dB, qp_is14, qw_dq, qw_dA, qw_dB
    type(c_ptr) :: qq_dq_i, qq_dA_i, qq_dB_i, qq_dq_j, qq_dA_j, qq_dB_j, qq_is14
  end type qnb_qtangent

  ! Q bonded terms (c_loc of the arrays qatom_load_fep fills, converted to
  ! 0-based atom/term numbers and per-term code rows)
  type, bind(c) :: qnb_qbonded
    integer(c_int32_t) :: nbond, nangle, ntorsion, nimproper
    integer(c_int32_t) :: nbond_lib, nangle_lib, ntorsion_lib, nimproper_lib
    type(c_ptr) :: bond_ij, bond_cod, bond_lib
    type(c_ptr) :: angle_ijk, angle_cod, angle_lib
    type(c_ptr) :: torsion_ijkl, torsion_cod, torsion_lib
    type(c_ptr) :: improper_ijkl, improper_cod, improper_lib
    integer(c_int32_t) :: nangle_coupl, ntorsion_coupl, nimproper_coupl
    type(c_ptr) :: angle_coupl, torsion_coupl, improper_coupl
    integer(c_int32_t) :: urey_bradley
  end type qnb_qbonded

  interface
    integer(c_int32_t) function qnb_abi_version() bind(c, name='qnb_abi_version')
      import :: c_int32_t
    end function
    type(c_ptr) function qnb_last_error() bind(c, name='qnb_last_error')
      import :: c_ptr
    end function
    integer(c_int32_t) function qnb_device_count() bind(c, name='qnb_device_count')
      import :: c_int32_t
    end function
    integer(c_int32_t) function qnb_create(sys, par, h) bind(c, name='qnb_create')
      import :: c_int32_t, c_ptr, qnb_system, qnb_params
      type(qnb_system), intent(in) :: sys
      type(qnb_params), intent(in) :: par
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int32_t) function qnb_destroy(h) bind(c, name='qnb_destroy')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: h
    end function
    ! make_pair_lists (src/nonbondene.f90:749-837)
    integer(c_int32_t) function qnb_make_pair_lists(h, x) bind(c, name='qnb_make_pair_lists')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
    end function
    ! pot_energy_nonbonds (src/potene.f90:320-380)
    integer(c_int32_t) function qnb_nonbonded(h, x, d, E, EQ) bind(c, name='qnb_nonbonded')
      import :: c_int32_t, c_ptr, c_double, qnb_energies, qnb_qenergies
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      type(qnb_energies), intent(inout) :: E
      type(qnb_qenergies), intent(inout) :: EQ(*)
    end function
    ! nonbond_qq + nonbond_qqp (src/nonbondene.f90:5013-5153)
    integer(c_int32_t) function qnb_qq(h, x, d, EQ) bind(c, name='qnb_qq')
      import :: c_int32_t, c_ptr, c_double, qnb_qenergies
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      type(qnb_qenergies), intent(inout) :: EQ(*)
    end function
    ! restrain_solvent + watpol (src/nonbondene.f90:6466-6746)
    integer(c_int32_t) function qnb_restraints(h, x, d, Tfree, istep, E) bind(c, name='qnb_restraints')
      import :: c_int32_t, c_ptr, c_double, qnb_energies
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      real(c_double), value :: Tfree
      integer(c_int32_t), value :: istep
      type(qnb_energies), intent(inout) :: E
    end function
    integer(c_int32_t) function qnb_eval_terms(h, mask, x, d, Tfree, istep, E, EQ) bind(c, name='qnb_eval_terms')
      import :: c_int32_t, c_ptr, c_double, qnb_energies, qnb_qenergies
      type(c_ptr), value :: h
      integer(c_int32_t), value :: mask
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      real(c_double), value :: Tfree
      integer(c_int32_t), value :: istep
      type(qnb_energies), intent(inout) :: E
      type(qnb_qenergies), intent(inout) :: EQ(*)
    end function
    integer(c_int32_t) function qnb_get_shell_state(h, st) bind(c, name='qnb_get_shell_state')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: st(*)
    end function
    integer(c_int32_t) function qnb_set_shell_state(h, st) bind(c, name='qnb_set_shell_state')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: st(*)
    end function
    ! adjoint of the per-state Q-atom energies over frames(natom*3, nframes);
    ! out(7, nqat, nstates, nframes)
    integer(c_int32_t) function qnb_adjoint(h, tan, nframes, frames, out) bind(c, name='qnb_adjoint')
      import :: c_int32_t, c_ptr, c_double, qnb_qtangent
      type(c_ptr), value :: h
      type(qnb_qtangent), intent(in) :: tan
      integer(c_int32_t), value :: nframes
      real(c_double), intent(in) :: frames(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int32_t) function qnb_set_qbonded(h, qb) bind(c, name='qnb_set_qbonded')
      import :: c_int32_t, c_ptr, qnb_qbonded
      type(c_ptr), value :: h
      type(qnb_qbonded), intent(in) :: qb
    end function
    ! pot_energy's Q bonded loop: adds into d(3, natom); eq(4, nstates)
    integer(c_int32_t) function qnb_eval_qbonded(h, x, d, eq) bind(c, name='qnb_eval_qbonded')
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      real(c_double), intent(out) :: eq(*)
    end function
  end interface

contains

  ! message of the last failed call, for die(msg) (src/qalloc.f90:609)
  function qnb_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: k
    msg = ''
    call c_f_pointer(qnb_last_error(), p, [256])
    do k = 1, 256
      if (p(k) == c_null_char) exit
      msg(k:k) = p(k)
    end do
  end function qnb_error_message

end module qnb_bind
