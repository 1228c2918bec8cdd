! negf.f90 -- the per-job interface of src/negf.f90:7 (Equilibrium_Density,
! Transmission_Coefficient, Local_Density_Of_State) plus batched forms that hand every
! job of a run to the GPU in one call.  The system (basis, V_reciprocal_all, geometry)
! is uploaded once and kept resident; it is re-uploaded only when the arguments change.
module negf
    use, intrinsic :: iso_c_binding
    use global
    use math_kernel, only: coefficient_simpson, value_simpson, fermi_func, gauss_legendre
    use negf_gpu_c
    implicit none
    private
    public Equilibrium_Density, Transmission_Coefficient, Local_Density_Of_State
    public Transmission_batch, Density_batch, LDOS_batch, contour_node, negf_release, negf_precision
    public negf_contour_rule, Bias_density_batch, negf_lead, LDOS_chunk

    type(c_ptr), save :: ctx = c_null_ptr
    complex*16, allocatable, save :: key_V(:, :, :)
    integer, allocatable, save :: key_nx(:), key_ny(:)
    real*8, save :: key_p(6) = 0.D0
    integer(c_int), save :: key_flags = -1
    integer(c_int), save :: negf_precision = NEGF_PREC_FP64
    ! NEGF_LEAD_DECIMATION or 0 (the reference's closed-form lead terms)
    integer(c_int), save :: negf_lead = 0
    ! 0: the reference's Simpson rule; 1: Gauss-Legendre nodes on the arc and the line
    integer, save :: negf_contour_rule = 0
contains
    ! Create (or reuse) the GPU context for this system.
    subroutine bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        type(t_parameters), intent(in) :: atomic
        real*8 :: p(6)
        integer(c_int), allocatable :: nx(:), ny(:)
        complex(c_double_complex), allocatable :: V(:, :, :)
        p = [atomic%LX, atomic%LY, atomic%LZ, atomic%V_L, atomic%V_R, atomic%ETA]
        if (c_associated(ctx)) then
            if (key_flags /= ior(negf_precision, negf_lead)) call negf_release()
        end if
        if (c_associated(ctx)) then
            if (all(shape(key_V) == shape(V_reciprocal_all)) .and. size(key_nx) == size(nx_grid)) then
                if (all(key_V == V_reciprocal_all) .and. all(key_nx == nx_grid) .and. &
                    all(key_ny == ny_grid) .and. all(key_p == p)) return
            end if
            call negf_release()
        end if
        allocate(nx(size(nx_grid)), ny(size(ny_grid)))
        nx = nx_grid
        ny = ny_grid
        allocate(V(size(V_reciprocal_all, 1), size(V_reciprocal_all, 2), size(V_reciprocal_all, 3)))
        V = V_reciprocal_all
        call negf_check(negf_ctx_create(ctx, size(nx), size(V, 3), nx, ny, atomic%LX, atomic%LY, &
                        atomic%LZ, atomic%V_L, atomic%V_R, atomic%ETA, V, (size(V, 1) - 1) / 2, &
                        (size(V, 2) - 1) / 2, ior(negf_precision, negf_lead)), "negf_ctx_create")
        key_V = V_reciprocal_all
        key_nx = nx_grid
        key_ny = ny_grid
        key_p = p
        key_flags = ior(negf_precision, negf_lead)
    end subroutine bind_system

    subroutine negf_release()
        integer(c_int) :: st
        if (c_associated(ctx)) st = negf_ctx_destroy(ctx)
        ctx = c_null_ptr
        if (allocated(key_V)) deallocate(key_V)
    end subroutine negf_release

    ! Node energy (without eta) and weight of contour point i_integral
    ! (semicircle then line; src/negf.f90:36-65, 86-111).
    subroutine contour_node(i_integral, N_circle, N_line, min_V, atomic, z, w)
        integer, intent(in) :: i_integral, N_circle, N_line
        real*8, intent(in) :: min_V
        type(t_parameters), intent(in) :: atomic
        complex*16, intent(out) :: z, w
        real*8 :: cl, cr, ll, lr, ec, er, th, gx, gw
        cl = min_V - 1.D0 / hartree
        cr = atomic%MU - atomic%GAP
        ll = atomic%MU - atomic%GAP
        lr = atomic%MU + atomic%GAP
        ec = (cr + cl) / 2
        er = (cr - cl) / 2
        if (negf_contour_rule == 1) then
            if (i_integral <= N_circle) then
                call gl_node(i_integral, N_circle, gx, gw)
                th = pi / 2 * (gx + 1.D0)
                z = ec + er * exp(cmplx(0.D0, th, kind=8))
                w = 2.D0 / pi * (pi / 2 * gw) * cmplx(0.D0, 1.D0, kind=8) * er * exp(cmplx(0.D0, th, kind=8))
            else
                call gl_node(i_integral - N_circle, N_line, gx, gw)
                z = cmplx((lr + ll) / 2 + (lr - ll) / 2 * gx, 0.D0, kind=8)
                w = -(2.D0 / pi * (lr - ll) / 2 * gw * fermi_func(dble(z), atomic%MU, atomic%TEMPERATURE))
            end if
            return
        end if
        if (i_integral <= N_circle) then
            th = value_simpson(i_integral, N_circle, 0.D0, pi)
            z = ec + er * exp(cmplx(0.D0, th, kind=8))
            w = 2.D0 / pi * coefficient_simpson(i_integral, N_circle, 0.D0, pi) * &
                cmplx(0.D0, 1.D0, kind=8) * er * exp(cmplx(0.D0, th, kind=8))
        else
            z = cmplx(value_simpson(i_integral - N_circle, N_line, ll, lr), 0.D0, kind=8)
            w = -(2.D0 / pi * coefficient_simpson(i_integral - N_circle, N_line, ll, lr) * &
                  fermi_func(dble(z), atomic%MU, atomic%TEMPERATURE))
        end if
    contains
        subroutine gl_node(i, n, gx, gw)
            integer, intent(in) :: i, n
            real*8, intent(out) :: gx, gw
            real*8, allocatable :: xs(:), ws(:)
            allocate(xs(n), ws(n))
            call gauss_legendre(n, xs, ws)
            gx = xs(i)
            gw = ws(i)
        end subroutine gl_node
    end subroutine contour_node

    subroutine Equilibrium_Density(i_job, V_reciprocal_all, nx_grid, ny_grid, N_circle, N_line, min_V, &
                                   atomic, kpoint, Density_Matrix, inverse_time)
        real*8, intent(in) :: min_V
        complex*16, intent(in), allocatable :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: i_job, N_circle, N_line, nx_grid(:), ny_grid(:)
        type(t_parameters), intent(in) :: atomic
        type(t_timer), intent(inout) :: inverse_time
        type(t_kpointmesh) :: kpoint(:)
        complex*16, intent(inout) :: Density_Matrix(:, :, :)
        integer :: i_integral, i_kpoint
        complex(c_double_complex) :: z(1), w(1)
        real(c_double) :: kx(1), ky(1), kw(1)
        i_integral = 1 + mod(i_job - 1, N_circle + N_line)
        i_kpoint = 1 + (i_job - 1) / (N_circle + N_line)
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        call contour_node(i_integral, N_circle, N_line, min_V, atomic, z(1), w(1))
        kx(1) = kpoint(i_kpoint)%kx
        ky(1) = kpoint(i_kpoint)%ky
        kw(1) = kpoint(i_kpoint)%weight
        call cpu_time(inverse_time%start)
        call negf_check(negf_density_batch(ctx, z, w, 1, kx, ky, kw, 1, Density_Matrix), "negf_density_batch")
        call cpu_time(inverse_time%end)
        inverse_time%sum = inverse_time%sum + inverse_time%end - inverse_time%start
    end subroutine Equilibrium_Density

    subroutine Transmission_Coefficient(i_job, V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, &
                                        Transmission, trans_time)
        complex*16, intent(in), allocatable :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: i_job, nx_grid(:), ny_grid(:)
        type(t_parameters), intent(in) :: atomic
        type(t_timer), intent(inout) :: trans_time
        type(t_kpointmesh) :: kpoint(:)
        type(t_transmission), intent(inout) :: Transmission(:)
        integer :: i_energy, i_kpoint
        real(c_double) :: E(1), kx(1), ky(1), kw(1), tau(1)
        i_energy = 1 + mod(i_job - 1, size(Transmission))
        i_kpoint = 1 + (i_job - 1) / size(Transmission)
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        E(1) = Transmission(i_energy)%energy
        kx(1) = kpoint(i_kpoint)%kx
        ky(1) = kpoint(i_kpoint)%ky
        kw(1) = kpoint(i_kpoint)%weight
        tau(1) = 0.D0
        call cpu_time(trans_time%start)
        call negf_check(negf_transmission_batch(ctx, E, 1, kx, ky, kw, 1, tau, c_null_ptr), &
                        "negf_transmission_batch")
        call cpu_time(trans_time%end)
        trans_time%sum = trans_time%sum + trans_time%end - trans_time%start
        Transmission(i_energy)%tau = Transmission(i_energy)%tau + tau(1)
    end subroutine Transmission_Coefficient

    subroutine Local_Density_Of_State(i_kpoint, i_energy, N_E, V_reciprocal_all, nx_grid, ny_grid, atomic, &
                                      kpoint, LDOS_Matrix, LDOS_inverse_time)
        complex*16, intent(in), allocatable :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: i_kpoint, i_energy, N_E, nx_grid(:), ny_grid(:)
        type(t_parameters), intent(in) :: atomic
        type(t_timer), intent(inout) :: LDOS_inverse_time
        type(t_kpointmesh) :: kpoint(:)
        complex*16, intent(inout) :: LDOS_Matrix(:, :, :)
        complex(c_double_complex), allocatable :: part(:, :, :)
        real(c_double) :: E(1), kx(1), ky(1), kw(1)
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        E(1) = atomic%LDOS_ENERGY_GRID(1) + &
               (atomic%LDOS_ENERGY_GRID(2) - atomic%LDOS_ENERGY_GRID(1)) * (i_energy - 1) / N_E
        kx(1) = kpoint(i_kpoint)%kx
        ky(1) = kpoint(i_kpoint)%ky
        kw(1) = kpoint(i_kpoint)%weight
        allocate(part(size(LDOS_Matrix, 1), size(LDOS_Matrix, 2), size(LDOS_Matrix, 3)))
        call cpu_time(LDOS_inverse_time%start)
        call negf_check(negf_ldos_batch(ctx, E, 1, kx, ky, kw, 1, part), "negf_ldos_batch")
        call cpu_time(LDOS_inverse_time%end)
        LDOS_inverse_time%sum = LDOS_inverse_time%sum + LDOS_inverse_time%end - LDOS_inverse_time%start
        LDOS_Matrix = LDOS_Matrix + part
    end subroutine Local_Density_Of_State

    ! Every (energy, k) job of the transmission loop in one launch (src/main.f90:384-398).
    subroutine Transmission_batch(V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, Transmission, seconds)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        type(t_parameters), intent(in) :: atomic
        type(t_kpointmesh), intent(in) :: kpoint(:)
        type(t_transmission), intent(inout) :: Transmission(:)
        real*8, intent(out) :: seconds
        real(c_double), allocatable :: E(:), kx(:), ky(:), kw(:), tau(:)
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        allocate(E(size(Transmission)), tau(size(Transmission)))
        allocate(kx(size(kpoint)), ky(size(kpoint)), kw(size(kpoint)))
        E = Transmission(:)%energy
        kx = kpoint(:)%kx
        ky = kpoint(:)%ky
        kw = kpoint(:)%weight
        tau = 0.D0
        call negf_check(negf_transmission_batch(ctx, E, size(E), kx, ky, kw, size(kx), tau, c_null_ptr), &
                        "negf_transmission_batch")
        seconds = negf_last_kernel_seconds(ctx)
        Transmission(:)%tau = Transmission(:)%tau + tau
    end subroutine Transmission_batch

    ! Every contour node and k point of the density loop in one launch (src/main.f90:272-288).
    subroutine Density_batch(V_reciprocal_all, nx_grid, ny_grid, N_circle, N_line, min_V, atomic, kpoint, &
                             Density_Matrix, seconds)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: nx_grid(:), ny_grid(:), N_circle, N_line
        real*8, intent(in) :: min_V
        type(t_parameters), intent(in) :: atomic
        type(t_kpointmesh), intent(in) :: kpoint(:)
        complex*16, intent(inout) :: Density_Matrix(:, :, :)
        real*8, intent(out) :: seconds
        complex(c_double_complex), allocatable :: z(:), w(:)
        real(c_double), allocatable :: kx(:), ky(:), kw(:)
        integer :: i
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        allocate(z(N_circle + N_line), w(N_circle + N_line))
        do i = 1, N_circle + N_line
            call contour_node(i, N_circle, N_line, min_V, atomic, z(i), w(i))
        end do
        allocate(kx(size(kpoint)), ky(size(kpoint)), kw(size(kpoint)))
        kx = kpoint(:)%kx
        ky = kpoint(:)%ky
        kw = kpoint(:)%weight
        call negf_check(negf_density_batch(ctx, z, w, size(z), kx, ky, kw, size(kx), Density_Matrix), &
                        "negf_density_batch")
        seconds = negf_last_kernel_seconds(ctx)
    end subroutine Density_batch

    ! Bias-window part of the non-equilibrium density (extension; the reference rejects
    ! VDS /= 0, src/main.f90:129-132), all window nodes and k points in one launch:
    !   Density_Matrix += i/pi sum_n q_n (f_L - f_R)(E_n) [G Gamma_L G^dagger]_kk (E_n + i eta)
    ! with mu_L = MU, mu_R = MU - VDS and E_n on [mu_R - GAP, mu_L + GAP] by the contour's
    ! rule (Simpson or Gauss-Legendre).  The i puts the term into Im(Density_Matrix), which
    ! the density reads (src/main.f90:347).
    subroutine Bias_density_batch(V_reciprocal_all, nx_grid, ny_grid, N_window, atomic, kpoint, &
                                  Density_Matrix, seconds)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        integer, intent(in) :: nx_grid(:), ny_grid(:), N_window
        type(t_parameters), intent(in) :: atomic
        type(t_kpointmesh), intent(in) :: kpoint(:)
        complex*16, intent(inout) :: Density_Matrix(:, :, :)
        real*8, intent(out) :: seconds
        real(c_double), allocatable :: E(:), kx(:), ky(:), kw(:), gx(:), gw(:)
        complex(c_double_complex), allocatable :: w(:)
        real*8 :: lo, hi, q, muL, muR
        integer :: i
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        muL = atomic%MU
        muR = atomic%MU - atomic%VDS
        lo = muR - atomic%GAP
        hi = muL + atomic%GAP
        allocate(E(N_window), w(N_window), gx(N_window), gw(N_window))
        if (negf_contour_rule == 1) call gauss_legendre(N_window, gx, gw)
        do i = 1, N_window
            if (negf_contour_rule == 1) then
                E(i) = (hi + lo) / 2 + (hi - lo) / 2 * gx(i)
                q = (hi - lo) / 2 * gw(i)
            else
                E(i) = value_simpson(i, N_window, lo, hi)
                q = coefficient_simpson(i, N_window, lo, hi)
            end if
            w(i) = cmplx(0.D0, q / pi * (fermi_func(E(i), muL, atomic%TEMPERATURE) - &
                         fermi_func(E(i), muR, atomic%TEMPERATURE)), kind=8)
        end do
        allocate(kx(size(kpoint)), ky(size(kpoint)), kw(size(kpoint)))
        kx = kpoint(:)%kx
        ky = kpoint(:)%ky
        kw = kpoint(:)%weight
        call negf_check(negf_neq_density_batch(ctx, E, w, N_window, kx, ky, kw, size(kx), Density_Matrix), &
                        "negf_neq_density_batch")
        seconds = negf_last_kernel_seconds(ctx)
    end subroutine Bias_density_batch

    ! LDOS_Matrix for energy indices ie0 .. ie0 + nc - 1 (all k points) in one launch;
    ! LDOS_Matrix(:, :, (ic - 1) N_z + 1 : ic N_z) holds energy ie0 + ic - 1.
    subroutine LDOS_chunk(ie0, nc, N_E, V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, LDOS_Matrix, seconds)
        integer, intent(in) :: ie0, nc, N_E, nx_grid(:), ny_grid(:)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        type(t_parameters), intent(in) :: atomic
        type(t_kpointmesh), intent(in) :: kpoint(:)
        complex*16, intent(out) :: LDOS_Matrix(:, :, :)
        real*8, intent(out) :: seconds
        real(c_double), allocatable :: E(:), kx(:), ky(:), kw(:)
        integer :: ic
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        allocate(E(nc))
        do ic = 1, nc
            ! src/negf.f90:203-204, as LDOS_batch
            E(ic) = atomic%LDOS_ENERGY_GRID(1) + &
                    (atomic%LDOS_ENERGY_GRID(2) - atomic%LDOS_ENERGY_GRID(1)) * (ie0 + ic - 2) / N_E
        end do
        allocate(kx(size(kpoint)), ky(size(kpoint)), kw(size(kpoint)))
        kx = kpoint(:)%kx
        ky = kpoint(:)%ky
        kw = kpoint(:)%weight
        call negf_check(negf_ldos_batch(ctx, E, nc, kx, ky, kw, size(kx), LDOS_Matrix), "negf_ldos_batch")
        seconds = negf_last_kernel_seconds(ctx)
    end subroutine LDOS_chunk

    ! LDOS_Matrix for one energy index (all k points in one launch).
    subroutine LDOS_batch(i_energy, N_E, V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, LDOS_Matrix, seconds)
        integer, intent(in) :: i_energy, N_E, nx_grid(:), ny_grid(:)
        complex*16, intent(in) :: V_reciprocal_all(:, :, :)
        type(t_parameters), intent(in) :: atomic
        type(t_kpointmesh), intent(in) :: kpoint(:)
        complex*16, intent(out) :: LDOS_Matrix(:, :, :)
        real*8, intent(out) :: seconds
        real(c_double) :: E(1)
        real(c_double), allocatable :: kx(:), ky(:), kw(:)
        call bind_system(V_reciprocal_all, nx_grid, ny_grid, atomic)
        E(1) = atomic%LDOS_ENERGY_GRID(1) + &
               (atomic%LDOS_ENERGY_GRID(2) - atomic%LDOS_ENERGY_GRID(1)) * (i_energy - 1) / N_E
        allocate(kx(size(kpoint)), ky(size(kpoint)), kw(size(kpoint)))
        kx = kpoint(:)%kx
        ky = kpoint(:)%ky
        kw = kpoint(:)%weight
        call negf_check(negf_ldos_batch(ctx, E, 1, kx, ky, kw, size(kx), LDOS_Matrix), "negf_ldos_batch")
        seconds = negf_last_kernel_seconds(ctx)
    end subroutine LDOS_batch
end module negf
