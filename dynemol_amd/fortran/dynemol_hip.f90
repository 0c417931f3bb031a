! ISO_C_BINDING interface to libdynemol_hip (include/dynemol_hip.h).
! Replaces the implicit-interface, trailing-underscore calls of the reference's
! GPU layer (GPU_Interface.cpp:74-91, Taylor_gpu.cpp:208-232): every routine
! returns an integer status (0 = success) and takes explicit kinds.
module dynemol_hip
    use iso_c_binding
    implicit none
    integer(c_int), parameter :: DH_OK = 0
    integer(c_int), parameter :: DH_MAT_S = 0, DH_MAT_H = 1, DH_MAT_L = 2, DH_MAT_LINV = 3, &
                                 DH_MAT_HORTH = 4, DH_MAT_HPRIME = 5
    interface
        function dh_last_error() bind(C, name="dh_last_error")
            import :: c_ptr
            type(c_ptr) :: dh_last_error
        end function
        integer(c_int) function dh_init(device, ctx) bind(C, name="dh_init")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr) :: ctx
        end function
        integer(c_int) function dh_finalize(ctx) bind(C, name="dh_finalize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dh_system_create(ctx, n_atoms, first_orb, atom_lmax, n_orb, orb_n, orb_l, orb_m, &
                orb_nzeta, orb_zeta, orb_coef, orb_ip, orb_kwh, orb_vshift, orb_frag, n_frag, n_wp, sys) &
                bind(C, name="dh_system_create")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n_atoms, n_orb, n_frag, n_wp
            integer(c_int), intent(in) :: first_orb(*), atom_lmax(*), orb_n(*), orb_l(*), orb_m(*), orb_nzeta(*), &
                                          orb_frag(*)
            real(c_double), intent(in) :: orb_zeta(*), orb_coef(*), orb_ip(*), orb_kwh(*), orb_vshift(*)
            type(c_ptr) :: sys
        end function
        integer(c_int) function dh_system_destroy(sys) bind(C, name="dh_system_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: sys
        end function
        integer(c_int) function dh_set_coords(sys, coords) bind(C, name="dh_set_coords")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            real(c_double), intent(in) :: coords(*)
        end function
        integer(c_int) function dh_build_operators(sys, cutoff, s_drop) bind(C, name="dh_build_operators")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            real(c_double), value :: cutoff, s_drop
        end function
        integer(c_int) function dh_load_operators(sys, S, h, ld) bind(C, name="dh_load_operators")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            real(c_double), intent(in) :: S(*), h(*)
            integer(c_int), value :: ld
        end function
        integer(c_int) function dh_orthogonalize(sys, want_hprime) bind(C, name="dh_orthogonalize")
            import :: c_int, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: want_hprime
        end function
        integer(c_int) function dh_get_matrix(sys, which, out, ld) bind(C, name="dh_get_matrix")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: which, ld
            real(c_double) :: out(*)
        end function
        integer(c_int) function dh_set_wavepacket(sys, wp, psi_bra, psi_ket) bind(C, name="dh_set_wavepacket")
            import :: c_int, c_double_complex, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp
            complex(c_double_complex), intent(in) :: psi_bra(*), psi_ket(*)
        end function
        integer(c_int) function dh_set_initial_state(sys, wp, ao_bra, ao_ket) bind(C, name="dh_set_initial_state")
            import :: c_int, c_double_complex, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp
            complex(c_double_complex), intent(in) :: ao_bra(*), ao_ket(*)
        end function
        integer(c_int) function dh_get_wavepacket(sys, wp, psi_bra, psi_ket) bind(C, name="dh_get_wavepacket")
            import :: c_int, c_double_complex, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp
            complex(c_double_complex) :: psi_bra(*), psi_ket(*)
        end function
        integer(c_int) function dh_propagate(sys, t_init, t_max, n_terms) bind(C, name="dh_propagate")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            real(c_double), value :: t_init, t_max
            integer(c_int) :: n_terms
        end function
        ! several nuclear steps and rebuilds in one call (Chebyshev_driver.f:102-169):
        ! coords(3, n_atoms, (n_steps - 1) / rebuild_every), pops(n_frag + 1, n_wp, n_steps),
        ! erg(n_wp, n_steps)
        integer(c_int) function dh_propagate_steps(sys, n_steps, dt, rebuild_every, coords, cutoff, s_drop, &
                                                   pops, erg, n_terms) bind(C, name="dh_propagate_steps")
            import :: c_int, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: n_steps, rebuild_every
            real(c_double), value :: dt, cutoff, s_drop
            real(c_double) :: coords(*), pops(*)
            complex(c_double_complex) :: erg(*)
            integer(c_int) :: n_terms
        end function
        integer(c_int) function dh_get_observables(sys, wp, ao_bra, pops, erg) bind(C, name="dh_get_observables")
            import :: c_int, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp
            complex(c_double_complex) :: ao_bra(*)
            real(c_double) :: pops(*), erg(2)
        end function
        ! eigenpairs il..iu (1-based) of h C = S C E; E and C may be c_null_ptr
        integer(c_int) function dh_eigenpairs(sys, il, iu, E, C, ldc) bind(C, name="dh_eigenpairs")
            import :: c_int, c_ptr
            type(c_ptr), value :: sys, E, C
            integer(c_int), value :: il, iu, ldc
        end function
        integer(c_int) function dh_set_eigenstate(sys, wp, k, energy) bind(C, name="dh_set_eigenstate")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp, k
            real(c_double) :: energy
        end function
        ! X: c_loc of real*8 X(ldx, N), or c_null_ptr to form X_ij from the system's basis
        integer(c_int) function dh_ehrenfest_kernel(sys, wp_el, wp_hl, X, ldx, K, ldk) bind(C, name="dh_ehrenfest_kernel")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: wp_el, wp_hl, ldx, ldk
            type(c_ptr), value :: X
            real(c_double) :: K(*)
        end function
        integer(c_int) function dh_propagation_elhl(ctx, N, S, h, H_out, AO_bra, AO_ket, PSI_bra, PSI_ket, &
                t_init, t_max, tau, save_tau) bind(C, name="dh_propagation_elhl")
            import :: c_int, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: N
            real(c_double), intent(in) :: S(*), h(*)
            real(c_double) :: H_out(*)
            complex(c_double_complex) :: AO_bra(*), AO_ket(*), PSI_bra(*), PSI_ket(*)
            real(c_double), intent(in) :: t_init, t_max
            real(c_double) :: tau, save_tau
        end function
    end interface
contains
    ! Message of the last failed call, as a Fortran string.
    function dh_error_message() result(msg)
        character(len=512) :: msg
        character(kind=c_char), pointer :: p(:)
        integer :: i
        call c_f_pointer(dh_last_error(), p, [512])
        msg = ' '
        do i = 1, 512
            if (p(i) == c_null_char) exit
            msg(i:i) = p(i)
        end do
    end function
end module dynemol_hip
