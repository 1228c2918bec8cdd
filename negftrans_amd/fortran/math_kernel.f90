! math_kernel.f90 -- the solver interface of src/math_kernel.f90:5, with the block
! tri-diagonal solve running on the GPU (negf_rgf_tri_solve) instead of LAPACK/BLAS.
module math_kernel
    use, intrinsic :: iso_c_binding
    use global
    use negf_gpu_c
    implicit none
    private
    public GreensFunction_tri_solver, coefficient_simpson, value_simpson, fermi_func, gauss_legendre
contains
    ! Blocks of the inverse of tridiag(I, G_inv_D(:,:,k), I): "D" diagonal, "F" first
    ! column, "L" last column (any combination), as src/math_kernel.f90:44-165; the
    ! components not requested are zero.
    subroutine GreensFunction_tri_solver(G_inv_D, G_Function, DFL)
        complex*16, intent(in), target :: G_inv_D(:, :, :)
        type(t_gfunc), intent(out) :: G_Function(:, :, :)
        character(*), intent(in) :: DFL
        complex(c_double_complex), allocatable, target :: Dg(:, :, :), Fc(:, :, :), Lc(:, :, :), Din(:, :, :)
        integer(c_int) :: mode, st
        integer :: N, N_z
        type(c_ptr) :: pd, pf, pl
        N = size(G_inv_D, 1)
        N_z = size(G_inv_D, 3)
        mode = 0
        if (scan(DFL, "Dd") /= 0) mode = ior(mode, NEGF_MODE_D)
        if (scan(DFL, "Ff") /= 0) mode = ior(mode, NEGF_MODE_F)
        if (scan(DFL, "Ll") /= 0) mode = ior(mode, NEGF_MODE_L)
        G_Function(:, :, :)%diagonal = (0.D0, 0.D0)
        G_Function(:, :, :)%first_column = (0.D0, 0.D0)
        G_Function(:, :, :)%last_column = (0.D0, 0.D0)
        if (mode == 0) return
        allocate(Din(N, N, N_z))
        Din = G_inv_D            ! contiguous copy (the dummy may be a strided section)
        pd = c_null_ptr
        pf = c_null_ptr
        pl = c_null_ptr
        if (iand(mode, NEGF_MODE_D) /= 0) then
            allocate(Dg(N, N, N_z))
            pd = c_loc(Dg)
        end if
        if (iand(mode, NEGF_MODE_F) /= 0) then
            allocate(Fc(N, N, N_z))
            pf = c_loc(Fc)
        end if
        if (iand(mode, NEGF_MODE_L) /= 0) then
            allocate(Lc(N, N, N_z))
            pl = c_loc(Lc)
        end if
        st = negf_rgf_tri_solve(N, N_z, Din, mode, pd, pf, pl)
        call negf_check(st, "negf_rgf_tri_solve")
        if (allocated(Dg)) G_Function(:, :, :)%diagonal = Dg
        if (allocated(Fc)) G_Function(:, :, :)%first_column = Fc
        if (allocated(Lc)) G_Function(:, :, :)%last_column = Lc
    end subroutine GreensFunction_tri_solver

    ! Composite Simpson weight c_i on [a, b] with N nodes (1-based i), the rule of
    ! src/math_kernel.f90:167-204: 1/3 rule for odd N; for even N the 1/3 rule up to node
    ! N-3 and the 3/8 rule on the last four nodes (N = 2 and N = 4 behave as there).
    function coefficient_simpson(i, N, a, b) result(coefficient)
        integer, intent(in) :: i, N
        real*8, intent(in) :: a, b
        real*8 :: coefficient
        real*8 :: inner
        inner = merge(4.D0 / 3, 2.D0 / 3, mod(i, 2) == 0)
        if (mod(N, 2) == 1) then
            coefficient = merge(1.D0 / 3, inner, i == 1 .or. i == N)
        else if (i == 1) then
            coefficient = 1.D0 / 3
        else if (i == N - 3) then
            coefficient = 1.D0 / 3 + 1.D0 * 3 / 8
        else if (i == N - 2 .or. i == N - 1) then
            coefficient = 3.D0 * 3 / 8
        else if (i == N) then
            coefficient = 1.D0 * 3 / 8
        else
            coefficient = inner
        end if
        coefficient = coefficient * (b - a) / (N - 1)
    end function coefficient_simpson

    ! Node x_i of the uniform grid on [a, b] (src/math_kernel.f90:206-216).
    function value_simpson(i, N, a, b) result(val)
        integer, intent(in) :: i, N
        real*8, intent(in) :: a, b
        real*8 :: val
        val = (i - 1) * (b - a) / (N - 1) + a
    end function value_simpson

    ! n-point Gauss-Legendre rule on [-1, 1], ascending nodes (extension, no reference
    ! counterpart): Newton iteration on P_n from the three-term recurrence, started at
    ! cos(pi (i - 1/4) / (n + 1/2)); w_i = 2 / ((1 - x_i^2) P_n'(x_i)^2).
    subroutine gauss_legendre(n, x, w)
        integer, intent(in) :: n
        real*8, intent(out) :: x(n), w(n)
        integer :: i, k, it
        real*8 :: z, p0, p1, p2, dp, dz
        do i = 1, n
            z = cos(pi * (i - 0.25D0) / (n + 0.5D0))
            do it = 1, 100
                p0 = 1.D0
                p1 = z
                do k = 2, n
                    p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
                    p0 = p1
                    p1 = p2
                end do
                dp = n * (z * p1 - p0) / (z * z - 1.D0)
                dz = p1 / dp
                z = z - dz
                if (abs(dz) < 1.D-15) exit
            end do
            p0 = 1.D0
            p1 = z
            do k = 2, n
                p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
                p0 = p1
                p1 = p2
            end do
            dp = n * (z * p1 - p0) / (z * z - 1.D0)
            x(n + 1 - i) = z
            w(n + 1 - i) = 2.D0 / ((1.D0 - z * z) * dp * dp)
        end do
    end subroutine gauss_legendre

    ! Fermi-Dirac occupation (src/math_kernel.f90:218-223), atomic units.
    function fermi_func(energy, mu, temperature) result(f_E)
        real*8, intent(in) :: energy, mu, temperature
        real*8 :: f_E
        f_E = 1.D0 / (exp((energy - mu) / (k_B * temperature)) + 1)
    end function fermi_func
end module math_kernel
