! negf_gpu_c.f90 -- iso_c_binding interfaces to libnegf_gpu.so (include/negf_gpu.h).
! complex*16 arrays are passed as complex(c_double_complex), layout-identical to the
! interleaved (re, im) doubles of the C side; all arrays keep Fortran order.
module negf_gpu_c
    use, intrinsic :: iso_c_binding
    implicit none
    public

    integer(c_int), parameter :: NEGF_PREC_FP64 = 0, NEGF_PREC_MIXED_FP32 = 1
    integer(c_int), parameter :: NEGF_FULL_WINDOW = 16, NEGF_LEAD_DECIMATION = 32
    integer(c_int), parameter :: NEGF_MODE_D = 1, NEGF_MODE_F = 2, NEGF_MODE_L = 4

    interface
        function negf_last_error() bind(C, name="negf_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function

        integer(c_int) function negf_device_count() bind(C, name="negf_device_count")
            import :: c_int
        end function

        integer(c_int) function negf_ctx_create(ctx, N, N_z, nx, ny, Lx, Ly, Lz, V_L, V_R, eta, &
                                                vrec, NGX, NGY, flags) bind(C, name="negf_ctx_create")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr) :: ctx
            integer(c_int), value :: N, N_z, NGX, NGY, flags
            integer(c_int) :: nx(*), ny(*)
            real(c_double), value :: Lx, Ly, Lz, V_L, V_R, eta
            complex(c_double_complex) :: vrec(*)
        end function

        integer(c_int) function negf_ctx_destroy(ctx) bind(C, name="negf_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function negf_transmission_batch(ctx, E, nE, kx, ky, kw, nk, tau, tau_jobs) &
                bind(C, name="negf_transmission_batch")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double) :: E(*), kx(*), ky(*), kw(*), tau(*)
            type(c_ptr), value :: tau_jobs
            integer(c_int), value :: nE, nk
        end function

        integer(c_int) function negf_rgf_tri_solve(N, N_z, G_inv_D, mode, diag, first, last) &
                bind(C, name="negf_rgf_tri_solve")
            import :: c_int, c_ptr, c_double_complex
            integer(c_int), value :: N, N_z, mode
            complex(c_double_complex) :: G_inv_D(*)
            type(c_ptr), value :: diag, first, last
        end function

        integer(c_int) function negf_density_batch(ctx, z, w, nnode, kx, ky, kw, nk, dm) &
                bind(C, name="negf_density_batch")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            complex(c_double_complex) :: z(*), w(*), dm(*)
            real(c_double) :: kx(*), ky(*), kw(*)
            integer(c_int), value :: nnode, nk
        end function

        integer(c_int) function negf_ldos_batch(ctx, E, nE, kx, ky, kw, nk, lm) bind(C, name="negf_ldos_batch")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double) :: E(*), kx(*), ky(*), kw(*)
            complex(c_double_complex) :: lm(*)
            integer(c_int), value :: nE, nk
        end function

        integer(c_int) function negf_pw_to_real(N, nx, ny, nmat, mats, N_x, N_y, out) &
                bind(C, name="negf_pw_to_real")
            import :: c_int, c_double_complex
            integer(c_int), value :: N, nmat, N_x, N_y
            integer(c_int) :: nx(*), ny(*)
            complex(c_double_complex) :: mats(*), out(*)
        end function

        integer(c_int) function negf_vhat_from_real(N_x, N_y, N_z, V_real, NGX, NGY, full_window, vrec) &
                bind(C, name="negf_vhat_from_real")
            import :: c_int, c_double, c_double_complex
            integer(c_int), value :: N_x, N_y, N_z, NGX, NGY, full_window
            real(c_double) :: V_real(*)
            complex(c_double_complex) :: vrec(*)
        end function

        integer(c_int) function negf_nonlocal_vhat_from_real(N_x, N_y, V_real, Rx, Ry, vrec) &
                bind(C, name="negf_nonlocal_vhat_from_real")
            import :: c_int, c_double_complex
            integer(c_int), value :: N_x, N_y, Rx, Ry
            complex(c_double_complex) :: V_real(*), vrec(*)
        end function

        integer(c_int) function negf_inverse_batch(N, count, A, inv) bind(C, name="negf_inverse_batch")
            import :: c_int, c_double_complex
            integer(c_int), value :: N, count
            complex(c_double_complex) :: A(*), inv(*)
        end function

        integer(c_int) function negf_matrix_product_batch(N, count, A, B, C) &
                bind(C, name="negf_matrix_product_batch")
            import :: c_int, c_double_complex
            integer(c_int), value :: N, count
            complex(c_double_complex) :: A(*), B(*), C(*)
        end function

        integer(c_int) function negf_hamiltonian(ctx, kx, ky, out) bind(C, name="negf_hamiltonian")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double), value :: kx, ky
            complex(c_double_complex) :: out(*)
        end function

        integer(c_int) function negf_e_minus_h(ctx, E_re, E_im, kx, ky, out) bind(C, name="negf_e_minus_h")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double), value :: E_re, E_im, kx, ky
            complex(c_double_complex) :: out(*)
        end function

        integer(c_int) function negf_hamiltonian_direct(N, N_z, nx, ny, Lx, Ly, kx, ky, vrec, NGX, NGY, out) &
                bind(C, name="negf_hamiltonian_direct")
            import :: c_int, c_double, c_double_complex
            integer(c_int), value :: N, N_z, NGX, NGY
            integer(c_int) :: nx(*), ny(*)
            real(c_double), value :: Lx, Ly, kx, ky
            complex(c_double_complex) :: vrec(*), out(*)
        end function

        integer(c_int) function negf_e_minus_h_from_h(N, N_z, H, E_re, E_im, nx, ny, Lx, Ly, Lz, kx, ky, &
                V_L, V_R, out) bind(C, name="negf_e_minus_h_from_h")
            import :: c_int, c_double, c_double_complex
            integer(c_int), value :: N, N_z
            complex(c_double_complex) :: H(*), out(*)
            integer(c_int) :: nx(*), ny(*)
            real(c_double), value :: E_re, E_im, Lx, Ly, Lz, kx, ky, V_L, V_R
        end function

        integer(c_int) function negf_transmission_trace(N, N_z, G, E, nx, ny, Lx, Ly, Lz, kx, ky, V_L, V_R, tau) &
                bind(C, name="negf_transmission_trace")
            import :: c_int, c_double, c_double_complex
            integer(c_int), value :: N, N_z
            complex(c_double_complex) :: G(*)
            real(c_double), value :: E, Lx, Ly, Lz, kx, ky, V_L, V_R
            integer(c_int) :: nx(*), ny(*)
            real(c_double) :: tau
        end function

        real(c_double) function negf_last_kernel_seconds(ctx) bind(C, name="negf_last_kernel_seconds")
            import :: c_double, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function negf_neq_density_batch(ctx, E, w, nE, kx, ky, kw, nk, dm) &
                bind(C, name="negf_neq_density_batch")
            import :: c_int, c_double, c_ptr, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double) :: E(*), kx(*), ky(*), kw(*)
            complex(c_double_complex) :: w(*), dm(*)
            integer(c_int), value :: nE, nk
        end function

        integer(c_int) function negf_last_fallbacks(ctx) bind(C, name="negf_last_fallbacks")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! Abort with the library's message on a negative status (bad argument / HIP error).
    subroutine negf_check(status, what)
        integer(c_int), intent(in) :: status
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n, i
        character(len=512) :: text
        if (status >= 0) return
        p = negf_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, msg, [n])
        text = ""
        do i = 1, min(n, 512)
            text(i:i) = msg(i)
        end do
        write(0, '(A, A, I0, A, A)') what, " failed, status ", status, ": ", trim(text)
        error stop 2
    end subroutine negf_check

end module negf_gpu_c
