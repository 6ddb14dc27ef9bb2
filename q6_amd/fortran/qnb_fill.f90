! qnb_fill.f90 -- fills the engine's qnb_system and qnb_params (include/qnb.h,
! module qnb_bind) from Qdyn6's own module state, for a Qdyn6 that calls the
! engine in place of pot_energy_nonbonds.
!
! Call qnb_fill_system once after precompute_interactions (src/qdyn.f90:153),
! when the topology, the FEP file and the precomputed pair tables are in
! place, then qnb_create(sys, par, h).  The routine makes 0-based copies of
! Qdyn6's arrays in the layouts the header gives (index shifts, flattened
! tables, logicals as bytes) and keeps them in this module until the next
! call or qnb_fill_release; qnb_create copies what it needs, so they may be
! released right after it.
!
! Sources of each field (reference files):
!   counts, charges, types, exclusions, charge groups   src/topo.f90:73-204
!   ljcod (allocated maxval(iac) square)                src/simprep.f90:4555-4560
!   pw/ww/qp/qq/qw precompute tables                    src/globals.f90:320-338,
!                                                       src/simprep.f90:3069-3317
!   Q-atoms, qconn                                      src/qatom.f90, src/globals.f90:427
!   cut-offs, sphere, SCAAS and polarisation shells     src/globals.f90, src/topo.f90:132
!   pairs inside a solvent molecule (nonbnd_solv_int)   src/simprep.f90:4004-4138
module qnb_fill
  use, intrinsic :: iso_c_binding
  use qnb_bind
  use MD
  use QATOM
  implicit none
  private
  public :: qnb_fill_system, qnb_fill_release

  real(c_double), allocatable, target, save :: f_crg(:), f_avdw(:), f_bvdw(:)
  real(c_double), allocatable, target, save :: f_pw(:, :), f_ww(:, :), f_qp(:, :), f_qq(:, :), f_qw(:, :)
  integer(c_int32_t), allocatable, target, save :: f_iac(:), f_iqatom(:), f_ljcod(:)
  integer(c_int32_t), allocatable, target, save :: f_first(:), f_last(:), f_iswitch(:), f_cgpatom(:)
  integer(c_int32_t), allocatable, target, save :: f_exlong(:), f_14long(:), f_iqseq(:), f_sint_ij(:)
  real(c_double), allocatable, target, save :: f_sint_par(:)
  integer(c_int8_t), allocatable, target, save :: f_excl(:), f_heavy(:), f_listex(:), f_list14(:)
  integer(c_int8_t), allocatable, target, save :: f_qconn3(:), f_qp_set(:), f_qq_set(:), f_qq_soft(:)
  ! a valid address for tables a system does not have (never read then)
  real(c_double), target, save :: f_dummy_d(1) = 0.0_c_double
  integer(c_int32_t), target, save :: f_dummy_i(1) = 0
  integer(c_int8_t), target, save :: f_dummy_b(1) = 0

contains

  integer(c_int8_t) function b8(l)
    logical, intent(in) :: l
    b8 = merge(1_c_int8_t, 0_c_int8_t, l)
  end function

  subroutine qnb_fill_release()
    if (allocated(f_crg)) deallocate(f_crg, f_avdw, f_bvdw, f_iac, f_iqatom, f_ljcod, f_excl, f_heavy, &
                                     f_first, f_last, f_iswitch, f_cgpatom)
    if (allocated(f_listex)) deallocate(f_listex, f_list14, f_exlong, f_14long)
    if (allocated(f_pw)) deallocate(f_pw)
    if (allocated(f_ww)) deallocate(f_ww)
    if (allocated(f_iqseq)) deallocate(f_iqseq)
    if (allocated(f_qconn3)) deallocate(f_qconn3)
    if (allocated(f_qp)) deallocate(f_qp, f_qp_set)
    if (allocated(f_qq)) deallocate(f_qq, f_qq_set, f_qq_soft)
    if (allocated(f_qw)) deallocate(f_qw)
    if (allocated(f_sint_ij)) deallocate(f_sint_ij, f_sint_par)
  end subroutine

  ! sys, par: filled for Qdyn6's current state; precision: QNB_PREC_FP64 or
  ! QNB_PREC_MIXED.  ref_quirks is set to 1 so that the engine builds the
  ! lists exactly as this Qdyn6 does (the nbpwlist test point).
  subroutine qnb_fill_system(sys, par, precision)
    type(qnb_system), intent(out) :: sys
    type(qnb_params), intent(out) :: par
    integer(c_int32_t), intent(in) :: precision
    integer :: i, k, j, s, t, n, nl, nt, nsol, sa, nq, ns, ng

    call qnb_fill_release()
    nsol = nat_solute
    sa = solv_atom
    nq = nqat
    ns = max(nstates, 1)
    ng = ncgp
    nt = natyps

    sys%natom = natom
    sys%nat_solute = nsol
    sys%nwat = nwat
    sys%solv_atom = sa
    sys%ncgp = ng
    sys%ncgp_solute = ncgp_solute
    sys%nqat = nq
    sys%nstates = nstates
    sys%ntypes = nt
    sys%ivdw_rule = ivdw_rule
    sys%solvent_spc = merge(1, 0, solvent_type == SOLVENT_SPC .and. ivdw_rule == VDW_GEOMETRIC)
    sys%nexlong = nexlong
    sys%n14long = n14long
    sys%any_atom_cutoff = merge(1, 0, iuse_switch_atom == 0)
    sys%el14_scale = el14_scale

    ! per atom
    allocate(f_crg(natom), f_iac(natom), f_iqatom(natom), f_excl(natom), f_heavy(natom))
    do i = 1, natom
      f_crg(i) = crg(i)
      f_iac(i) = int(iac(i)) - 1
      f_iqatom(i) = int(iqatom(i))
      f_excl(i) = b8(excl(i))
      f_heavy(i) = b8(heavy(i))
    end do

    ! charge groups: ranges into cgpatom, switch atom
    n = cgp(ng)%last
    allocate(f_first(ng), f_last(ng), f_iswitch(ng), f_cgpatom(n))
    do i = 1, ng
      f_first(i) = cgp(i)%first - 1
      f_last(i) = cgp(i)%last - 1
      f_iswitch(i) = cgp(i)%iswitch - 1
    end do
    do i = 1, n
      f_cgpatom(i) = cgpatom(i) - 1
    end do

    ! atom types: (t, k) at t*3 + k; ljcod (ti, tj) at ti + tj*ntypes, types
    ! above maxval(iac) (no atoms) keep the default code 1
    allocate(f_avdw(3*nt), f_bvdw(3*nt), f_ljcod(nt*nt))
    do t = 1, nt
      do k = 1, 3
        f_avdw(3*(t-1) + k) = iaclib(t)%avdw(k)
        f_bvdw(3*(t-1) + k) = iaclib(t)%bvdw(k)
      end do
    end do
    f_ljcod(:) = 1
    nl = size(ljcod, 1)
    do j = 1, nl
      do i = 1, nl
        f_ljcod(i + nt*(j-1)) = ljcod(i, j)
      end do
    end do

    ! solute exclusion windows (i, i+k) at i*25 + k-1, long lists as 0-based pairs
    allocate(f_listex(max(1, max_nbr_range*nsol)), f_list14(max(1, max_nbr_range*nsol)))
    allocate(f_exlong(max(1, 2*nexlong)), f_14long(max(1, 2*n14long)))
    f_listex(:) = 0
    f_list14(:) = 0
    f_exlong(:) = 0
    f_14long(:) = 0
    do i = 1, nsol
      do k = 1, max_nbr_range
        f_listex(k + max_nbr_range*(i-1)) = b8(listex(k, i))
        f_list14(k + max_nbr_range*(i-1)) = b8(list14(k, i))
      end do
    end do
    do i = 1, nexlong
      f_exlong(2*i-1) = listexlong(1, i) - 1
      f_exlong(2*i) = listexlong(2, i) - 1
    end do
    do i = 1, n14long
      f_14long(2*i-1) = list14long(1, i) - 1
      f_14long(2*i) = list14long(2, i) - 1
    end do

    ! solute-solvent (i, j) at i + j*nat_solute from pw_precomp(i - pw_low, j);
    ! solvent-solvent (la, ka) at la + ka*solv_atom
    if (nwat > 0 .and. nsol > 0 .and. allocated(pw_precomp)) then
      allocate(f_pw(nsol*sa, 3))
      f_pw(:, :) = 0.0_c_double
      do j = 1, sa
        do i = 1, nsol
          if (i - pw_low >= 1 .and. i - pw_low <= size(pw_precomp, 1)) then
            n = i + nsol*(j-1)
            f_pw(n, 1) = pw_precomp(i - pw_low, j)%vdWA
            f_pw(n, 2) = pw_precomp(i - pw_low, j)%vdWB
            f_pw(n, 3) = pw_precomp(i - pw_low, j)%elec
          end if
        end do
      end do
    end if
    if (nwat > 0) then
      allocate(f_ww(sa*sa, 3))
      do k = 1, sa
        do i = 1, sa
          f_ww(i + sa*(k-1), 1) = ww_precomp(i, k)%vdWA
          f_ww(i + sa*(k-1), 2) = ww_precomp(i, k)%vdWB
          f_ww(i + sa*(k-1), 3) = ww_precomp(i, k)%elec
        end do
      end do
    end if

    ! Q-atoms: sequence, bonded-neighbour flags, per-state tables
    if (nq > 0) then
      allocate(f_iqseq(nq))
      do i = 1, nq
        f_iqseq(i) = iqseq(i) - 1
      end do
      if (nsol > 0) then
        allocate(f_qconn3(nsol), f_qp(nsol*nq*ns, 4), f_qp_set(nsol*nq*ns))
        do i = 1, nsol
          f_qconn3(i) = b8(any(qconn(:, i, :) <= 3))
        end do
        do s = 1, ns
          do k = 1, nq
            do i = 1, nsol
              n = i + nsol*((k-1) + nq*(s-1))
              f_qp(n, 1) = qp_precomp(i, k, s)%vdWA
              f_qp(n, 2) = qp_precomp(i, k, s)%vdWB
              f_qp(n, 3) = qp_precomp(i, k, s)%elec
              f_qp(n, 4) = qp_precomp(i, k, s)%score
              f_qp_set(n) = b8(qp_precomp(i, k, s)%set)
            end do
          end do
        end do
      end if
      allocate(f_qq(nq*nq*ns, 4), f_qq_set(nq*nq*ns), f_qq_soft(nq*nq*ns))
      do s = 1, ns
        do j = 1, nq
          do i = 1, nq
            n = i + nq*((j-1) + nq*(s-1))
            f_qq(n, 1) = qq_precomp(i, j, s)%vdWA
            f_qq(n, 2) = qq_precomp(i, j, s)%vdWB
            f_qq(n, 3) = qq_precomp(i, j, s)%elec
            f_qq(n, 4) = qq_precomp(i, j, s)%score
            f_qq_set(n) = b8(qq_precomp(i, j, s)%set)
            f_qq_soft(n) = b8(qq_precomp(i, j, s)%soft)
          end do
        end do
      end do
      if (nwat > 0) then
        allocate(f_qw(nq*sa*ns, 4))
        do s = 1, ns
          do j = 1, sa
            do i = 1, nq
              n = i + nq*((j-1) + sa*(s-1))
              f_qw(n, 1) = qw_precomp(i, j, s)%vdWA
              f_qw(n, 2) = qw_precomp(i, j, s)%vdWB
              f_qw(n, 3) = qw_precomp(i, j, s)%elec
              f_qw(n, 4) = qw_precomp(i, j, s)%score
            end do
          end do
        end do
      end if
    end if

    sys%crg = c_loc(f_crg)
    sys%iac = c_loc(f_iac)
    sys%iqatom = c_loc(f_iqatom)
    sys%excl = c_loc(f_excl)
    sys%heavy = c_loc(f_heavy)
    sys%cgp_first = c_loc(f_first)
    sys%cgp_last = c_loc(f_last)
    sys%cgp_iswitch = c_loc(f_iswitch)
    sys%cgpatom = c_loc(f_cgpatom)
    sys%avdw = c_loc(f_avdw)
    sys%bvdw = c_loc(f_bvdw)
    sys%ljcod = c_loc(f_ljcod)
    sys%listex = c_loc(f_listex)
    sys%list14 = c_loc(f_list14)
    sys%listexlong = c_loc(f_exlong)
    sys%list14long = c_loc(f_14long)
    if (allocated(f_pw)) then
      sys%pw_A = c_loc(f_pw(1, 1))
      sys%pw_B = c_loc(f_pw(1, 2))
      sys%pw_el = c_loc(f_pw(1, 3))
    else
      sys%pw_A = c_loc(f_dummy_d)
      sys%pw_B = c_loc(f_dummy_d)
      sys%pw_el = c_loc(f_dummy_d)
    end if
    if (allocated(f_ww)) then
      sys%ww_A = c_loc(f_ww(1, 1))
      sys%ww_B = c_loc(f_ww(1, 2))
      sys%ww_el = c_loc(f_ww(1, 3))
    else
      sys%ww_A = c_loc(f_dummy_d)
      sys%ww_B = c_loc(f_dummy_d)
      sys%ww_el = c_loc(f_dummy_d)
    end if
    sys%iqseq = c_loc(f_dummy_i)
    sys%qconn3 = c_loc(f_dummy_b)
    sys%qp_A = c_loc(f_dummy_d)
    sys%qp_B = c_loc(f_dummy_d)
    sys%qp_el = c_loc(f_dummy_d)
    sys%qp_sc = c_loc(f_dummy_d)
    sys%qp_set = c_loc(f_dummy_b)
    sys%qq_A = c_loc(f_dummy_d)
    sys%qq_B = c_loc(f_dummy_d)
    sys%qq_el = c_loc(f_dummy_d)
    sys%qq_sc = c_loc(f_dummy_d)
    sys%qq_set = c_loc(f_dummy_b)
    sys%qq_soft = c_loc(f_dummy_b)
    sys%qw_A = c_loc(f_dummy_d)
    sys%qw_B = c_loc(f_dummy_d)
    sys%qw_el = c_loc(f_dummy_d)
    sys%qw_sc = c_loc(f_dummy_d)
    if (allocated(f_iqseq)) sys%iqseq = c_loc(f_iqseq)
    if (allocated(f_qconn3)) then
      sys%qconn3 = c_loc(f_qconn3)
      sys%qp_A = c_loc(f_qp(1, 1))
      sys%qp_B = c_loc(f_qp(1, 2))
      sys%qp_el = c_loc(f_qp(1, 3))
      sys%qp_sc = c_loc(f_qp(1, 4))
      sys%qp_set = c_loc(f_qp_set)
    end if
    if (allocated(f_qq)) then
      sys%qq_A = c_loc(f_qq(1, 1))
      sys%qq_B = c_loc(f_qq(1, 2))
      sys%qq_el = c_loc(f_qq(1, 3))
      sys%qq_sc = c_loc(f_qq(1, 4))
      sys%qq_set = c_loc(f_qq_set)
      sys%qq_soft = c_loc(f_qq_soft)
    end if
    if (allocated(f_qw)) then
      sys%qw_A = c_loc(f_qw(1, 1))
      sys%qw_B = c_loc(f_qw(1, 2))
      sys%qw_el = c_loc(f_qw(1, 3))
      sys%qw_sc = c_loc(f_qw(1, 4))
    end if

    ! pairs inside a solvent molecule: listed only when the solvent has
    ! torsions (nonbond_solvent_internal runs only then); offsets 0-based
    sys%n_solv_int = 0
    sys%solv_int_ij = c_loc(f_dummy_i)
    sys%solv_int_par = c_loc(f_dummy_d)
    if (nwat > 0 .and. ntors > ntors_solute .and. allocated(nonbnd_solv_int)) then
      n = num_solv_int
      allocate(f_sint_ij(2*n), f_sint_par(3*n))
      do i = 1, n
        f_sint_ij(2*i-1) = nonbnd_solv_int(i)%i - 1
        f_sint_ij(2*i) = nonbnd_solv_int(i)%j - 1
        f_sint_par(3*i-2) = nonbnd_solv_int(i)%vdWA
        f_sint_par(3*i-1) = nonbnd_solv_int(i)%vdWB
        f_sint_par(3*i) = nonbnd_solv_int(i)%elec
      end do
      sys%n_solv_int = n
      sys%solv_int_ij = c_loc(f_sint_ij)
      sys%solv_int_par = c_loc(f_sint_par)
    end if

    ! run parameters
    par%Rcpp = Rcpp
    par%Rcpw = Rcpw
    par%Rcww = Rcww
    par%Rcq = Rcq
    par%RcLRF = RcLRF
    par%use_LRF = merge(1, 0, use_LRF)
    par%precision = precision
    par%ref_quirks = 1
    par%xpcent = (/xpcent%x, xpcent%y, xpcent%z/)
    par%xwcent = (/xwcent%x, xwcent%y, xwcent%z/)
    par%rexcl_o = rexcl_o
    par%lambda(:) = 0.0_c_double
    if (nstates > 0) par%lambda(1:nstates) = EQ(1:nstates)%lambda
    par%rwat = rwat
    par%fk_wsphere = fk_wsphere
    par%Dwmz = Dwmz
    par%awmz = awmz
    par%Boltz = Boltz
    par%wpol_restr = merge(1, 0, wpol_restr)
    par%nwpolr_shell = 0
    par%fkwpol = fkwpol
    par%itdis = itdis
    par%wshell_rout(:) = 0.0_c_double
    par%wshell_dr(:) = 0.0_c_double
    par%wshell_cstb(:) = 0.0_c_double
    par%wshell_theta_corr(:) = 0.0_c_double
    if (nwat > 0 .and. nwpolr_shell > 0) then
      n = nwpolr_shell
      if (n > QNB_MAX_SHELLS) call die('qnb_fill: more polarisation shells than the engine holds')
      par%nwpolr_shell = n
      par%wshell_rout(1:n) = wshell(1:n)%rout
      par%wshell_dr(1:n) = wshell(1:n)%dr
      par%wshell_cstb(1:n) = wshell(1:n)%cstb
      par%wshell_theta_corr(1:n) = wshell(1:n)%theta_corr
    end if
  end subroutine

end module qnb_fill
