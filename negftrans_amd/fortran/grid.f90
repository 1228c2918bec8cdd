! grid.f90 -- the basis / block / transform interface of src/grid.f90:7-10.  Basis and
! k mesh are host set-up; every transform and block builder runs on the GPU.
module grid
    use, intrinsic :: iso_c_binding
    use global
    use negf_gpu_c
    implicit none
    private
    public LocalPotential_RealToPlanewave, NonLocalPotential_RealToPlanewave, PlaneWaveBasis_construction_findsize, &
        Greensfunction_PlanewaveToReal, PlaneWaveBasis_construction, Hamiltonian_construction, &
        E_minus_H_construction, Kpoint_mesh_construction, Transmission_solver
contains
    ! V_reciprocal(p, q) = (1/(N_x N_y)) sum V(x, y) exp(-2 pi i (p x/N_x + q y/N_y)) over the
    ! whole window of V_reciprocal (bounds -Nr/2..Nr/2), src/grid.f90:12-69.
    subroutine LocalPotential_RealToPlanewave(V_real, V_reciprocal)
        real*8, intent(in) :: V_real(:, :)
        complex*16, allocatable, intent(inout) :: V_reciprocal(:, :)
        real(c_double), allocatable :: V(:, :)
        complex(c_double_complex), allocatable :: out(:, :)
        integer :: N_x, N_y, ngx, ngy
        N_x = size(V_real, 1)
        N_y = size(V_real, 2)
        if (size(V_reciprocal, 1) - 1 > N_x .or. size(V_reciprocal, 2) - 1 > N_y) then
            write(0, *) "ERROR: 'The size of V_reciprocal minus one' should not be larger than 'the size of V_real'"
            error stop 1
        end if
        if (-lbound(V_reciprocal, 1) /= ubound(V_reciprocal, 1) .or. &
            -lbound(V_reciprocal, 2) /= ubound(V_reciprocal, 2)) then
            write(0, *) "ERROR: The index boundary of V_reciprocal must be symmetric about zero"
            error stop 1
        end if
        ngx = ubound(V_reciprocal, 1)
        ngy = ubound(V_reciprocal, 2)
        allocate(V(N_x, N_y), out(-ngx:ngx, -ngy:ngy))
        V = V_real
        call negf_check(negf_vhat_from_real(N_x, N_y, 1, V, ngx, ngy, 1, out), "negf_vhat_from_real")
        V_reciprocal = out
    end subroutine LocalPotential_RealToPlanewave

    ! V_reciprocal(p, q, p', q') = (N_x N_y)^-2 sum V(x, y, x', y') exp(-2 pi i (p x/N_x + q y/N_y
    ! - p' x'/N_x - q' y'/N_y)) over the whole symmetric window, src/grid.f90:137-234 (the
    ! reference's validation messages kept).
    subroutine NonLocalPotential_RealToPlanewave(V_real, V_reciprocal)
        complex*16, intent(in) :: V_real(:, :, :, :)
        complex*16, allocatable, intent(inout) :: V_reciprocal(:, :, :, :)
        complex(c_double_complex), allocatable :: V(:, :, :, :), out(:, :, :, :)
        integer :: N_x, N_y, rx, ry
        N_x = size(V_real, 1)
        N_y = size(V_real, 2)
        if (size(V_reciprocal, 1) - 1 > N_x .or. size(V_reciprocal, 2) - 1 > N_y) then
            write(0, *) "ERROR: 'The size of V_reciprocal minus one' should not be larger than 'the size of V_real'"
            error stop 1
        end if
        if (-lbound(V_reciprocal, 1) /= ubound(V_reciprocal, 1) .or. &
            -lbound(V_reciprocal, 2) /= ubound(V_reciprocal, 2)) then
            write(0, *) "ERROR: The index boundary of V_reciprocal must be symmetric about zero"
            error stop 1
        end if
        rx = ubound(V_reciprocal, 1)
        ry = ubound(V_reciprocal, 2)
        allocate(V(N_x, N_y, N_x, N_y), out(-rx:rx, -ry:ry, -rx:rx, -ry:ry))
        V = V_real
        call negf_check(negf_nonlocal_vhat_from_real(N_x, N_y, V, rx, ry, out), "negf_nonlocal_vhat_from_real")
        V_reciprocal = out
    end subroutine NonLocalPotential_RealToPlanewave

    ! Real-space diagonal of one plane-wave block (src/grid.f90:71-135).
    subroutine Greensfunction_PlanewaveToReal(subGreenFunc, nx_grid, ny_grid, subGreenDiag)
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        complex*16, intent(in) :: subGreenFunc(:, :)
        complex*16, intent(out) :: subGreenDiag(:, :)
        integer(c_int), allocatable :: nx(:), ny(:)
        complex(c_double_complex), allocatable :: M(:, :), out(:, :)
        allocate(nx(size(nx_grid)), ny(size(ny_grid)), M(size(subGreenFunc, 1), size(subGreenFunc, 2)))
        allocate(out(size(subGreenDiag, 1), size(subGreenDiag, 2)))
        nx = nx_grid
        ny = ny_grid
        M = subGreenFunc
        call negf_check(negf_pw_to_real(size(nx), nx, ny, 1, M, size(out, 1), size(out, 2), out), &
                        "negf_pw_to_real")
        subGreenDiag = out
    end subroutine Greensfunction_PlanewaveToReal

    ! H(i, j, k) = V(nx_i - nx_j, ny_i - ny_j, k) + delta_ij |G_i + k|^2 / 2 (src/grid.f90:236-262).
    subroutine Hamiltonian_construction(nx_grid, ny_grid, Lx, Ly, kx, ky, V_reciprocal, Hamiltonian)
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        complex*16, allocatable, intent(in) :: V_reciprocal(:, :, :)
        real*8, intent(in) :: Lx, Ly, kx, ky
        complex*16, intent(out) :: Hamiltonian(:, :, :)
        integer(c_int), allocatable :: nx(:), ny(:)
        complex(c_double_complex), allocatable :: V(:, :, :), H(:, :, :)
        allocate(nx(size(nx_grid)), ny(size(ny_grid)))
        nx = nx_grid
        ny = ny_grid
        allocate(V(size(V_reciprocal, 1), size(V_reciprocal, 2), size(V_reciprocal, 3)))
        V = V_reciprocal
        allocate(H(size(Hamiltonian, 1), size(Hamiltonian, 2), size(Hamiltonian, 3)))
        call negf_check(negf_hamiltonian_direct(size(nx), size(H, 3), nx, ny, Lx, Ly, kx, ky, V, &
                        ubound(V_reciprocal, 1), ubound(V_reciprocal, 2), H), "negf_hamiltonian_direct")
        Hamiltonian = H
    end subroutine Hamiltonian_construction

    ! D_k = 2 Delta^2 (energy I - H_k) - 2 I plus the closed-form lead terms
    ! exp(i q Delta) on the first and last slice (src/grid.f90:264-333).
    subroutine E_minus_H_construction(Hamiltonian, energy, nx_grid, ny_grid, Lx, Ly, Lz, &
                                      kx, ky, V_L, V_R, E_minus_H)
        complex*16, intent(in) :: Hamiltonian(:, :, :)
        complex*16, intent(in) :: energy
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        real*8, intent(in) :: Lx, Ly, Lz, kx, ky, V_L, V_R
        complex*16, intent(out) :: E_minus_H(:, :, :)
        integer(c_int), allocatable :: nx(:), ny(:)
        complex(c_double_complex), allocatable :: H(:, :, :), D(:, :, :)
        allocate(nx(size(nx_grid)), ny(size(ny_grid)))
        nx = nx_grid
        ny = ny_grid
        allocate(H(size(Hamiltonian, 1), size(Hamiltonian, 2), size(Hamiltonian, 3)))
        allocate(D(size(Hamiltonian, 1), size(Hamiltonian, 2), size(Hamiltonian, 3)))
        H = Hamiltonian
        call negf_check(negf_e_minus_h_from_h(size(H, 1), size(H, 3), H, dble(energy), aimag(energy), &
                        nx, ny, Lx, Ly, Lz, kx, ky, V_L, V_R, D), "negf_e_minus_h_from_h")
        E_minus_H = D
    end subroutine E_minus_H_construction

    ! tau = sum_ij gamma_L(i) gamma_R(j) |G_{N_z,1}(i, j)|^2 at real energy (src/grid.f90:335-385).
    subroutine Transmission_solver(G_Function_first_column, energy, nx_grid, ny_grid, Lx, Ly, Lz, &
                                   kx, ky, V_L, V_R, tau)
        complex*16, intent(in) :: G_Function_first_column(:, :, :)
        integer, intent(in) :: nx_grid(:), ny_grid(:)
        real*8, intent(in) :: energy, Lx, Ly, Lz, kx, ky, V_L, V_R
        real*8, intent(out) :: tau
        integer(c_int), allocatable :: nx(:), ny(:)
        complex(c_double_complex), allocatable :: G(:, :)
        real(c_double) :: t
        integer :: N_z
        N_z = size(G_Function_first_column, 3)
        allocate(nx(size(nx_grid)), ny(size(ny_grid)))
        nx = nx_grid
        ny = ny_grid
        allocate(G(size(G_Function_first_column, 1), size(G_Function_first_column, 2)))
        G = G_Function_first_column(:, :, N_z)
        call negf_check(negf_transmission_trace(size(G, 1), N_z, G, energy, nx, ny, Lx, Ly, Lz, kx, ky, &
                        V_L, V_R, t), "negf_transmission_trace")
        tau = t
    end subroutine Transmission_solver

    ! Basis size: lattice points (i, j) with ((2 pi i/Lx)^2 + (2 pi j/Ly)^2)/2 <= ENCUT.
    function PlaneWaveBasis_construction_findsize(ENCUT, Lx, Ly) result(N)
        real*8, intent(in) :: ENCUT, Lx, Ly
        integer :: N
        integer :: i, j, mx, my
        mx = ceiling(Lx / pi * sqrt(ENCUT / 2))
        my = ceiling(Ly / pi * sqrt(ENCUT / 2))
        N = 0
        do j = -my, my
            do i = -mx, mx
                if (((2 * i * pi / Lx) ** 2 + (2 * j * pi / Ly) ** 2) / 2 <= ENCUT) N = N + 1
            end do
        end do
    end function PlaneWaveBasis_construction_findsize

    ! The basis in the reference's order: n_y outer ascending, n_x inner ascending
    ! (src/grid.f90:410-431).
    subroutine PlaneWaveBasis_construction(ENCUT, Lx, Ly, nx_grid, ny_grid)
        real*8, intent(in) :: ENCUT, Lx, Ly
        integer, intent(out) :: nx_grid(:), ny_grid(:)
        integer :: i, j, mx, my, n
        mx = ceiling(Lx / pi * sqrt(ENCUT / 2))
        my = ceiling(Ly / pi * sqrt(ENCUT / 2))
        n = 0
        do j = -my, my
            do i = -mx, mx
                if (((2 * i * pi / Lx) ** 2 + (2 * j * pi / Ly) ** 2) / 2 <= ENCUT) then
                    n = n + 1
                    nx_grid(n) = i
                    ny_grid(n) = j
                end if
            end do
        end do
    end subroutine PlaneWaveBasis_construction

    ! Monkhorst-Pack mesh, x index fastest, uniform weights (src/grid.f90:433-448).
    subroutine Kpoint_mesh_construction(NKX, NKY, Lx, Ly, kpoint)
        integer, intent(in) :: NKX, NKY
        real*8, intent(in) :: Lx, Ly
        type(t_kpointmesh), intent(out) :: kpoint(:)
        integer :: i, j, n
        do j = 1, NKY
            do i = 1, NKX
                n = i + (j - 1) * NKX
                kpoint(n)%kx = (i - (1 + NKX) / 2.D0) / NKX * (2 * pi / Lx)
                kpoint(n)%ky = (j - (1 + NKY) / 2.D0) / NKY * (2 * pi / Ly)
                kpoint(n)%weight = 1.D0 / (NKX * NKY)
            end do
        end do
    end subroutine Kpoint_mesh_construction
end module grid
