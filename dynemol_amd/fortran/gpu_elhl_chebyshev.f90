! Drop-in for the reference's gpu_ElHl_Chebyshev_m (ElHl_Chebyshev_GPU.f:28-318)
! and the CPU twin ElHl_Chebyshev_m (ElHl_Chebyshev.f:24-291): the same two
! subroutines and argument meaning, over device-resident state in
! libdynemol_hip.  The electron (column 1) and the hole (column 2) are two
! wavepackets of one device system instead of two MPI ranks.
!
! The record types carry only the fields the propagator path reads
! (types_EHT.f:8-45, 114-166); the initial state (the FMO electron/hole
! orbitals, FMO.f:45-139) arrives in AO_bra/AO_ket from the caller.
module dh_types_m
    implicit none
    type structure
        integer :: atoms
        real*8, allocatable :: coord(:, :)        ! (atoms, 3), Angstrom
        integer, allocatable :: fragment(:)       ! fragment index per atom (1-based)
    end type
    type STO_basis
        integer :: n, l, m, Nzeta, atom, fragment
        real*8 :: IP, k_WH, V_shift
        real*8 :: zeta(2), coef(2)
    end type
    type f_time
        integer :: n_fragments
        real*8, allocatable :: dyn(:, :, :)       ! (it, 0:n_fragments+1, n_part)
    end type
end module dh_types_m

module gpu_ElHl_Chebyshev_m
    use iso_c_binding
    use dynemol_hip
    use dh_types_m
    implicit none
    private
    public :: gpu_preprocess_ElHl_Chebyshev, gpu_ElHl_Chebyshev, gpu_ElHl_finalize, QM_wp_erg, gpu_EhrenfestKernel

    ! a fresh start or a restart, told apart by the argument list, as in the
    ! reference's generic (ElHl_Chebyshev_GPU.f:45-48)
    interface gpu_preprocess_ElHl_Chebyshev
        module procedure preprocess_fresh, preprocess_from_restart
    end interface

    real*8, parameter :: frame_step = 1.d0      ! parameters_m frame_step for slice_Cheb
    type(c_ptr), save :: ctx = c_null_ptr, sysh = c_null_ptr
    logical, save :: first_call = .true.
    real*8, allocatable, save :: last_coord(:, :)
    complex*16, save :: QM_wp_erg(2)            ! Unit_Cell%QM_wp_erg (ElHl_Chebyshev.f:365-369)

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        if (rc /= DH_OK) then
            print *, 'dynemol_hip: ', what, ' failed: ', trim(dh_error_message())
            stop 1
        end if
    end subroutine

    ! The device system of the electron and the hole, its S and h built from the
    ! coordinates (Overlap_Matrix + Build_Huckel of both preprocess routines)
    subroutine create_system(system, basis, nf)
        type(structure), intent(in) :: system
        type(STO_basis), intent(in) :: basis(:)
        integer, intent(in) :: nf
        integer :: N, i, a
        integer(c_int), allocatable :: first_orb(:), lmax(:), on(:), ol(:), om(:), onz(:), ofr(:)
        real(c_double), allocatable :: oz(:), oc(:), oip(:), okwh(:), ovs(:), xyz(:)

        N = size(basis)
        allocate(first_orb(system%atoms + 1), lmax(system%atoms), on(N), ol(N), om(N), onz(N), ofr(N))
        allocate(oz(2 * N), oc(2 * N), oip(N), okwh(N), ovs(N), xyz(3 * system%atoms))
        lmax = 0
        first_orb = N
        do i = N, 1, -1          ! BasisPointer: first orbital of each atom (0-based)
            a = basis(i)%atom
            first_orb(a) = i - 1
            lmax(a) = max(lmax(a), basis(i)%l)
        end do
        first_orb(system%atoms + 1) = N
        do i = 1, N
            on(i) = basis(i)%n; ol(i) = basis(i)%l; om(i) = basis(i)%m; onz(i) = basis(i)%Nzeta
            ofr(i) = basis(i)%fragment - 1
            oz(2 * i - 1:2 * i) = basis(i)%zeta; oc(2 * i - 1:2 * i) = basis(i)%coef
            oip(i) = basis(i)%IP; okwh(i) = basis(i)%k_WH; ovs(i) = basis(i)%V_shift
        end do
        if (.not. c_associated(ctx)) call check(dh_init(0_c_int, ctx), 'dh_init')
        call check(dh_system_create(ctx, int(system%atoms, c_int), first_orb, lmax, int(N, c_int), on, ol, om, onz, &
                   oz, oc, oip, okwh, ovs, ofr, int(nf, c_int), 2_c_int, sysh), 'dh_system_create')
        do a = 1, system%atoms
            xyz(3 * a - 2:3 * a) = system%coord(a, :)
        end do
        call check(dh_set_coords(sysh, xyz), 'dh_set_coords')
        call check(dh_build_operators(sysh, 12.d0, 0.d0), 'dh_build_operators')
        if (allocated(last_coord)) deallocate(last_coord)
        allocate(last_coord, source=system%coord)
    end subroutine create_system

    ! preprocess_ElHl_Chebyshev (ElHl_Chebyshev.f:50-143) / gpu_preprocess (ElHl_Chebyshev_GPU.f:55-160)
    subroutine preprocess_fresh(system, basis, AO_bra, AO_ket, Dual_bra, Dual_ket, QDyn, it)
        type(structure), intent(in) :: system
        type(STO_basis), intent(in) :: basis(:)
        complex*16, intent(inout) :: AO_bra(:, :), AO_ket(:, :)
        complex*16, intent(out) :: Dual_bra(:, :), Dual_ket(:, :)
        type(f_time), intent(inout) :: QDyn
        integer, intent(in) :: it
        integer :: N, p, nf
        complex*16, allocatable :: psi_bra(:), psi_ket(:), ao(:)
        real*8 :: pops(QDyn%n_fragments + 1), erg(2)

        N = size(basis)
        nf = QDyn%n_fragments
        call create_system(system, basis, nf)
        allocate(psi_bra(N), psi_ket(N), ao(N))
        do p = 1, 2
            call check(dh_set_initial_state(sysh, int(p - 1, c_int), AO_bra(:, p), AO_ket(:, p)), 'dh_set_initial_state')
        end do
        do p = 1, 2
            call check(dh_get_wavepacket(sysh, int(p - 1, c_int), psi_bra, psi_ket), 'dh_get_wavepacket')
            call check(dh_get_observables(sysh, int(p - 1, c_int), ao, pops, erg), 'dh_get_observables')
            ! DUAL_bra = AO_bra, DUAL_ket = S*AO_ket (ElHl_Chebyshev.f:110-115)
            Dual_bra(:, p) = AO_bra(:, p)
            Dual_ket(:, p) = psi_bra
            QDyn%dyn(it, 0, p) = 0.d0
            QDyn%dyn(it, 1:nf + 1, p) = pops
            QM_wp_erg(p) = cmplx(erg(1), erg(2), kind=8)
        end do
        first_call = .true.
    end subroutine preprocess_fresh

    ! preprocess_from_restart (ElHl_Chebyshev.f:377-433): the states of a
    ! Restart_copy.dat snapshot (Restart_State_Cheb, backup.f:405-445) become
    ! Psi_bra = DUAL_ket and Psi_ket = AO_ket; QM_wp_erg follows from them.
    subroutine preprocess_from_restart(system, basis, DUAL_ket, AO_bra, AO_ket)
        type(structure), intent(in) :: system
        type(STO_basis), intent(in) :: basis(:)
        complex*16, intent(in) :: DUAL_ket(:, :), AO_bra(:, :), AO_ket(:, :)
        integer :: N, p, nf
        integer(c_int) :: nt
        complex*16, allocatable :: ao(:)
        real*8, allocatable :: pops(:)
        real*8 :: erg(2)

        N = size(basis)
        nf = maxval(basis%fragment)
        call create_system(system, basis, nf)
        allocate(ao(N), pops(nf + 1))
        do p = 1, 2
            call check(dh_set_wavepacket(sysh, int(p - 1, c_int), DUAL_ket(:, p), AO_ket(:, p)), 'dh_set_wavepacket')
        end do
        call check(dh_propagate(sysh, 0.d0, 0.d0, nt), 'dh_propagate')  ! zero step: the outputs of the restored states
        do p = 1, 2
            call check(dh_get_observables(sysh, int(p - 1, c_int), ao, pops, erg), 'dh_get_observables')
            QM_wp_erg(p) = cmplx(erg(1), erg(2), kind=8)
        end do
        first_call = .true.
    end subroutine preprocess_from_restart

    ! ElHl_Chebyshev (ElHl_Chebyshev.f:148-291) / gpu_ElHl_Chebyshev (ElHl_Chebyshev_GPU.f:168-318)
    subroutine gpu_ElHl_Chebyshev(system, basis, AO_bra, AO_ket, Dual_bra, Dual_ket, QDyn, t, delta_t, it)
        type(structure), intent(in) :: system
        type(STO_basis), intent(in) :: basis(:)
        complex*16, intent(inout) :: AO_bra(:, :), AO_ket(:, :), Dual_bra(:, :), Dual_ket(:, :)
        type(f_time), intent(inout) :: QDyn
        real*8, intent(inout) :: t
        real*8, intent(in) :: delta_t
        integer, intent(in) :: it
        integer :: N, p, a, nf
        integer(c_int) :: nterms
        real*8 :: t_init, t_max, pops(QDyn%n_fragments + 1), erg(2)
        real(c_double), allocatable :: xyz(:)
        complex*16, allocatable :: psi_bra(:), psi_ket(:), ao(:)

        N = size(basis)
        nf = QDyn%n_fragments
        t_init = t
        t_max = delta_t * frame_step * (it - 1)
        ! S and h are rebuilt when the nuclei moved (ElHl_Chebyshev.f:194-210)
        if (any(system%coord /= last_coord)) then
            allocate(xyz(3 * system%atoms))
            do a = 1, system%atoms
                xyz(3 * a - 2:3 * a) = system%coord(a, :)
            end do
            call check(dh_set_coords(sysh, xyz), 'dh_set_coords')
            call check(dh_build_operators(sysh, 12.d0, 0.d0), 'dh_build_operators')
            last_coord = system%coord
        end if
        call check(dh_propagate(sysh, t_init, t_max, nterms), 'dh_propagate')
        t = t_init + delta_t * frame_step
        allocate(psi_bra(N), psi_ket(N), ao(N))
        do p = 1, 2
            call check(dh_get_wavepacket(sysh, int(p - 1, c_int), psi_bra, psi_ket), 'dh_get_wavepacket')
            call check(dh_get_observables(sysh, int(p - 1, c_int), ao, pops, erg), 'dh_get_observables')
            Dual_bra(:, p) = conjg(psi_ket)       ! ElHl_Chebyshev.f:270-271
            Dual_ket(:, p) = psi_bra
            AO_bra(:, p) = conjg(ao)              ! :274-276 (S^-1 Psi_bra, conjugated)
            AO_ket(:, p) = psi_ket
            QDyn%dyn(it, 0, p) = t                ! Populations (data_output.f:87-147)
            QDyn%dyn(it, 1:nf + 1, p) = pops
            QM_wp_erg(p) = cmplx(erg(1), erg(2), kind=8)
        end do
        first_call = .false.
    end subroutine gpu_ElHl_Chebyshev

    ! Diabatic-Ehrenfest kernel matrix of the current electron/hole state, for
    ! Diabatic_Ehrenfest (diabatic-Ehren.f:100-116): replaces calculate_rho,
    ! A_ad_nd = (rho_eh + transpose(rho_eh))/two, Huckel_stuff and
    ! EhrenfestKernel_gpu(N, H_prime, A_ad_nd, X_ij, Kernel).  H', A and X stay on
    ! the device; only Kernel(N, N) comes back.
    subroutine gpu_EhrenfestKernel(Kernel)
        real*8, intent(out) :: Kernel(:, :)
        call check(dh_ehrenfest_kernel(sysh, 0_c_int, 1_c_int, c_null_ptr, int(size(Kernel, 1), c_int), Kernel, &
                                       int(size(Kernel, 1), c_int)), 'dh_ehrenfest_kernel')
    end subroutine

    subroutine gpu_ElHl_finalize()
        integer(c_int) :: rc
        if (c_associated(sysh)) rc = dh_system_destroy(sysh)
        if (c_associated(ctx)) rc = dh_finalize(ctx)
        sysh = c_null_ptr
        ctx = c_null_ptr
        if (allocated(last_coord)) deallocate(last_coord)
    end subroutine
end module gpu_ElHl_Chebyshev_m
