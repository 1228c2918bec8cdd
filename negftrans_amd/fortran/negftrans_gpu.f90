! negftrans_gpu.f90 -- NEGFtrans-compatible driver on one MI355X.
!
! Reads the same INPUT namelist and POTENTIAL file as the reference program
! (src/main.f90:52-54, 153-160; doc/Manual.tex sections "Input Files") and writes the
! same OUTPUT, DENSITY, TRANSMISSION and LDOS files (src/main.f90:562-601).  The energy
! loops are not walked job by job: each phase hands all of its (energy or contour node,
! k point) jobs to the GPU in one batched call (module negf).  There is one process and
! one GPU, so the MPI distribution of the reference has no counterpart here.
!
! Extension keys of the namelist (defaults keep reference behaviour):
!   PRECISION_MODE = 0 (complex128) | 1 (mixed: FP32 sweep + FP64 refinement, transmission)
!   CONTOUR_QUADRATURE = 'SIMPSON' (the reference's rule) | 'GAUSS-LEGENDRE' (density contour)
!   FULL_WINDOW    = .false. (copy only |p|<=NGX/2, |q|<=NGY/2 of V-hat, as the reference)
!   DUMP_FULL      = .false. (also write *_FULL files with 17 significant digits)
!   LEAD_SELF_ENERGY = 'CLOSED-FORM' (the reference's exp(i q Delta)) | 'DECIMATION' (lattice
!                    lead, decimated on the GPU; a V = 0 device transmits 1 per open channel)
!   BIAS_DENSITY   = .false. (.true. accepts VDS > 0: the density is the equilibrium contour
!                    at mu_R = MU - VDS plus the bias window [mu_R - GAP, MU + GAP] with
!                    G Gamma_L G^dagger; the reference rejects VDS /= 0)
program negftrans_gpu
    use, intrinsic :: iso_c_binding
    use global
    use grid
    use negf
    use tools
    use negf_gpu_c
    implicit none

    logical :: if_density = .false., if_transmission = .false., if_LDOS = .false.
    real*8 :: V_L = 0.D0, V_R = 0.D0, MU = 5.568D0, ETA = 1.D-5, TEMPERATURE = 20.D0
    real*8 :: LX = 0.D0, LY = 0.D0, LZ = 0.D0, ENCUT = 100.D0, GAP = 6.D0, VDS = 0.D0
    real*8 :: TRANSMISSION_GRID(3) = 0.D0, LDOS_ENERGY_GRID(3) = 0.D0
    integer :: LDOS_GRID(3) = -10, NGX = 0, NGY = 0, NKX = 0, NKY = 0
    integer :: N_circle = 40, N_line_per_eV = 500
    integer :: PRECISION_MODE = 0
    character(len=32) :: CONTOUR_QUADRATURE = 'SIMPSON', LEAD_SELF_ENERGY = 'CLOSED-FORM'
    logical :: FULL_WINDOW = .false., DUMP_FULL = .false., BIAS_DENSITY = .false.
    namelist /INPUT/ V_L, V_R, MU, ETA, TEMPERATURE, LX, LY, LZ, NGX, NGY, ENCUT, &
        N_circle, N_line_per_eV, GAP, NKX, NKY, VDS, TRANSMISSION_GRID, if_density, if_transmission, &
        if_LDOS, LDOS_ENERGY_GRID, LDOS_GRID, PRECISION_MODE, FULL_WINDOW, DUMP_FULL, CONTOUR_QUADRATURE, &
        BIAS_DENSITY, LEAD_SELF_ENERGY

    type(t_parameters) :: atomic, atomic_eq
    type(t_kpointmesh), allocatable :: kpoint(:)
    type(t_transmission), allocatable :: Transmission(:)
    real*8, allocatable :: V_real(:, :, :), Density(:, :, :), LDOS(:, :, :, :)
    complex*16, allocatable :: V_reciprocal_all(:, :, :), Density_Matrix(:, :, :), LDOS_Matrix(:, :, :)
    complex(c_double_complex), allocatable :: realspace(:, :, :)
    integer, allocatable :: nx_grid(:), ny_grid(:)
    integer(c_int), allocatable :: cnx(:), cny(:)
    integer :: nchunk, nc, ie0, ic
    integer :: N_x, N_y, N_z, N, N_E, N_line, N_window, i, j, k, ii, jj, kk, ie, io, bound
    real*8 :: rescale_factor, gpu_s, t_total0, t_total1, t_rtop, t_dens, t_ptor, t_trans, t_ldos, t_ldos_ptor
    real*8 :: t0, t1, gpu_s2
    logical :: bad

    call wall(t_total0)
    t_rtop = 0; t_dens = 0; t_ptor = 0; t_trans = 0; t_ldos = 0; t_ldos_ptor = 0
    open(unit=16, file="OUTPUT")
    call banner(16)
    write(16, *) "Read in INPUT..."
    open(unit=15, file="INPUT", status="old", iostat=io)
    if (io /= 0) then
        write(16, *) "ERROR: cannot open INPUT"
        call stop_run()
    end if
    read(15, nml=INPUT)
    close(15)

    ! NGX/NGY automatic values (src/main.f90:94-100; the NGX test uses the LY bound there)
    bound = ceiling((LY / a_0) / pi * sqrt(ENCUT / hartree / 2) * 2)
    if (NGX <= bound) NGX = ceiling((LX / a_0) / pi * sqrt(ENCUT / hartree / 2) * 2) + 1
    if (NGY <= bound) NGY = bound + 1

    write(16, *) "Checking Parameters..."
    bad = .false.
    if (LX == 0.D0 .or. LY == 0.D0 .or. LZ == 0.D0) then
        write(16, *) "ERROR: LX, LY, LZ must be specified in INPUT"
        bad = .true.
    end if
    if (NKX == 0 .or. NKY == 0) then
        write(16, *) "ERROR: NKX, NKY must be specified in INPUT"
        bad = .true.
    end if
    if (.not. (if_density .or. if_transmission .or. if_LDOS)) then
        write(16, *) "ERROR: At least one of (if_xxxxx) be .true.   Nothing to do..."
        bad = .true.
    end if
    if (if_transmission .and. TRANSMISSION_GRID(3) == 0.D0) then
        write(16, *) "ERROR: TRANSMISSION_GRID must be specified for (if_transmission = .true.)"
        bad = .true.
    end if
    if (if_LDOS .and. LDOS_ENERGY_GRID(3) == 0.D0) then
        write(16, *) "ERROR: LDOS_ENERGY_GRID must be specified for (if_LDOS = .true.)"
        bad = .true.
    end if
    if (if_LDOS .and. any(LDOS_GRID < -1)) then
        write(16, *) "ERROR: LDOS_GRID must be specified for (if_LDOS = .true.)"
        bad = .true.
    end if
    if (VDS /= 0.D0 .and. .not. BIAS_DENSITY) then
        write(16, *) "ERROR: VDS must be zero in this version, please upgrade to higher version."
        bad = .true.
    end if
    if (BIAS_DENSITY .and. VDS < 0.D0) then
        write(16, *) "ERROR: BIAS_DENSITY needs VDS >= 0 (the left lead at the higher Fermi level)"
        bad = .true.
    end if
    if (bad) call stop_run()
    write(16, *) "-success-"
    write(16, *) "The Parameters are:"
    write(16, nml=INPUT)

    write(16, *) "----------"
    write(16, *) "Read in POTENTIAL..."
    open(unit=17, file="POTENTIAL", status="old", iostat=io)
    if (io /= 0) then
        write(16, *) "ERROR: cannot open POTENTIAL"
        call stop_run()
    end if
    read(17, *) N_x, N_y, N_z
    allocate(V_real(N_x, N_y, N_z))
    read(17, *) (((V_real(i, j, k), i = 1, N_x), j = 1, N_y), k = 1, N_z)
    close(17)
    if (mod(N_x, 2) == 1 .or. mod(N_y, 2) == 1) then
        write(16, *) "ERROR: The number of grid point in x & y direction must be even number"
        call stop_run()
    end if
    write(16, *) "-success-"
    write(16, *) "The size of POTENTIAL is:", N_x, N_y, N_z
    write(16, *) "----------"

    ! unit conversion (src/main.f90:182-205)
    atomic%V_L = V_L / hartree
    atomic%V_R = V_R / hartree
    atomic%MU = MU / hartree
    atomic%ETA = ETA
    atomic%TEMPERATURE = TEMPERATURE
    atomic%LX = LX / a_0
    atomic%LY = LY / a_0
    atomic%LZ = LZ / a_0
    atomic%ENCUT = ENCUT / hartree
    atomic%GAP = GAP * k_B * atomic%TEMPERATURE
    atomic%VDS = VDS / hartree
    atomic%TRANSMISSION_GRID = TRANSMISSION_GRID(1:2) / hartree
    atomic%LDOS_ENERGY_GRID = LDOS_ENERGY_GRID(1:2) / hartree
    V_real = V_real / hartree

    allocate(kpoint(NKX * NKY))
    call Kpoint_mesh_construction(NKX, NKY, atomic%LX, atomic%LY, kpoint)
    write(16, *) "Number of K-point grid is:", size(kpoint)
    write(16, *) "   kx          ky           weight"
    write(16, '(3G12.3)') (kpoint(i)%kx, kpoint(i)%ky, kpoint(i)%weight, i = 1, size(kpoint))
    write(16, *) "----------"

    N = PlaneWaveBasis_construction_findsize(atomic%ENCUT, atomic%LX, atomic%LY)
    allocate(nx_grid(N), ny_grid(N), cnx(N), cny(N))
    call PlaneWaveBasis_construction(atomic%ENCUT, atomic%LX, atomic%LY, nx_grid, ny_grid)
    cnx = nx_grid
    cny = ny_grid
    write(16, *) "Layout the Grid..."
    write(16, *) "Grid size in x, y direction is: N =", N
    write(16, *) "Grid size in z direction is: N_z =", N_z
    write(16, *) "The 'Inverse Matrix' time is proportional to (N^3 N_z)"
    write(16, *) "----------"
    write(16, *) "Transform Potential to plane wave basis..."
    negf_precision = PRECISION_MODE
    select case (trim(adjustl(CONTOUR_QUADRATURE)))
    case ('SIMPSON', 'simpson')
        negf_contour_rule = 0
    case ('GAUSS-LEGENDRE', 'gauss-legendre')
        negf_contour_rule = 1
    case default
        write(16, *) "ERROR: CONTOUR_QUADRATURE must be SIMPSON or GAUSS-LEGENDRE"
        call stop_run()
    end select
    select case (trim(adjustl(LEAD_SELF_ENERGY)))
    case ('CLOSED-FORM', 'closed-form')
        negf_lead = 0
    case ('DECIMATION', 'decimation')
        negf_lead = NEGF_LEAD_DECIMATION
    case default
        write(16, *) "ERROR: LEAD_SELF_ENERGY must be CLOSED-FORM or DECIMATION"
        call stop_run()
    end select

    ! V_reciprocal_all for every slice on the GPU, window as the reference copies it
    call wall(t0)
    allocate(V_reciprocal_all(-NGX:NGX, -NGY:NGY, N_z))
    call negf_check(negf_vhat_from_real(N_x, N_y, N_z, V_real, NGX, NGY, merge(1, 0, FULL_WINDOW), &
                    V_reciprocal_all), "negf_vhat_from_real")
    call wall(t1)
    t_rtop = t1 - t0
    write(16, *) "DONE"
    write(16, *) "----------"

    if (if_density) then
        N_line = nint(N_line_per_eV * ((2 * atomic%GAP) * hartree))
        write(16, *) "Density calculation start..."
        write(16, *) "Number of energy points for integration:", N_circle + N_line
        write(16, *) "  hemi-circle integral:", N_circle
        write(16, *) "  line integral:", N_line
        allocate(Density(N_x, N_y, N_z), Density_Matrix(N, N, N_z), realspace(N_x, N_y, N_z))
        Density_Matrix = (0.D0, 0.D0)
        call wall(t0)
        if (BIAS_DENSITY .and. atomic%VDS > 0.D0) then
            ! equilibrium part up to the lower level mu_R = MU - VDS, then the window
            atomic_eq = atomic
            atomic_eq%MU = atomic%MU - atomic%VDS
            N_window = max(3, nint(N_line_per_eV * ((atomic%VDS + 2 * atomic%GAP) * hartree)))
            write(16, *) "  bias window integral:", N_window
            call Density_batch(V_reciprocal_all, nx_grid, ny_grid, N_circle, N_line, minval(V_real), &
                               atomic_eq, kpoint, Density_Matrix, gpu_s)
            call Bias_density_batch(V_reciprocal_all, nx_grid, ny_grid, N_window, atomic, kpoint, &
                                    Density_Matrix, gpu_s2)
            gpu_s = gpu_s + gpu_s2
        else
            call Density_batch(V_reciprocal_all, nx_grid, ny_grid, N_circle, N_line, minval(V_real), atomic, &
                               kpoint, Density_Matrix, gpu_s)
        end if
        call wall(t1)
        t_dens = t1 - t0
        write(16, *) "Transform Green's Function from plane wave basis to real space:"
        call wall(t0)
        call negf_check(negf_pw_to_real(N, cnx, cny, N_z, Density_Matrix, N_x, N_y, realspace), "negf_pw_to_real")
        Density = aimag(realspace) * (dble(N_z) / (LX * LY * LZ))       ! Angstrom^-3
        call wall(t1)
        t_ptor = t1 - t0
        write(16, *) "Density calculation DONE"
        write(16, *) "----------"
        deallocate(realspace)
    end if

    if (if_transmission) then
        N_E = nint(TRANSMISSION_GRID(3))
        allocate(Transmission(N_E))
        do i = 1, N_E
            Transmission(i)%energy = atomic%TRANSMISSION_GRID(1) + &
                (atomic%TRANSMISSION_GRID(2) - atomic%TRANSMISSION_GRID(1)) * (i - 1) / N_E
        end do
        Transmission(:)%tau = 0.D0
        write(16, *) "Transmission calculation start..."
        write(16, *) "Number of energy grid points:", N_E
        call wall(t0)
        call Transmission_batch(V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, Transmission, gpu_s)
        call wall(t1)
        t_trans = t1 - t0
        write(16, *) "Transmission calculation DONE"
        write(16, *) "----------"
    end if

    if (if_LDOS) then
        N_E = nint(LDOS_ENERGY_GRID(3))
        ! energies per GPU call: as many as fit 2 GB of host blocks and real-space slices
        nchunk = max(1, min(N_E, int(2.0D9 / (16.D0 * (dble(N) * N * N_z + dble(N_x) * N_y * N_z)))))
        allocate(LDOS_Matrix(N, N, N_z * nchunk), realspace(N_x, N_y, N_z * nchunk))
        allocate(LDOS(LDOS_size(LDOS_GRID(1), N_x), LDOS_size(LDOS_GRID(2), N_y), &
                      LDOS_size(LDOS_GRID(3), N_z), N_E))
        LDOS = 0.D0
        write(16, *) "LDOS calculation start..."
        write(16, *) "Number of energy grid points:", N_E
        do ie0 = 1, N_E, nchunk
            nc = min(nchunk, N_E - ie0 + 1)
            call wall(t0)
            call LDOS_chunk(ie0, nc, N_E, V_reciprocal_all, nx_grid, ny_grid, atomic, kpoint, LDOS_Matrix, gpu_s)
            call wall(t1)
            t_ldos = t_ldos + (t1 - t0)
            call negf_check(negf_pw_to_real(N, cnx, cny, N_z * nc, LDOS_Matrix, N_x, N_y, realspace), &
                            "negf_pw_to_real")
            do ic = 1, nc
                ie = ie0 + ic - 1
                do k = 1, N_z
                    do j = 1, N_y
                        do i = 1, N_x
                            rescale_factor = dble(N_z) / (LX * LY * LZ) / hartree
                            call LDOS_grid_converter(i, LDOS_GRID(1), LX / N_x, ii, rescale_factor)
                            call LDOS_grid_converter(j, LDOS_GRID(2), LY / N_y, jj, rescale_factor)
                            call LDOS_grid_converter(k, LDOS_GRID(3), LZ / N_z, kk, rescale_factor)
                            LDOS(ii, jj, kk, ie) = LDOS(ii, jj, kk, ie) + &
                                                   aimag(realspace(i, j, (ic - 1) * N_z + k)) * rescale_factor
                        end do
                    end do
                end do
                write(16, '(A2, I6, A2, I6)') "=>", ie, " /", N_E
            end do
            call wall(t0)
            t_ldos_ptor = t_ldos_ptor + (t0 - t1)
        end do
        write(16, *) "LDOS calculation DONE"
        write(16, *) "----------"
    end if

    write(16, *) ""
    write(16, *) "===== ALL NEGF calculation DONE ====="
    write(16, *) ""
    if (if_density) then
        write(16, *) "Total charge is:", sum(Density) * (LX * LY * LZ) / (N_x * N_y * N_z)
        write(16, *) "Writing DENSITY..."
        open(unit=18, file="DENSITY")
        write(18, '(3I5)') N_x, N_y, N_z
        write(18, '(5G17.8)') (((Density(i, j, k), i = 1, N_x), j = 1, N_y), k = 1, N_z)
        close(18)
        if (DUMP_FULL) then
            open(unit=18, file="DENSITY_FULL")
            write(18, '(3I6)') N_x, N_y, N_z
            write(18, '(ES25.17)') (((Density(i, j, k), i = 1, N_x), j = 1, N_y), k = 1, N_z)
            close(18)
        end if
    end if
    if (if_transmission) then
        write(16, *) "Writing TRANSMISSION..."
        open(unit=19, file="TRANSMISSION")
        write(19, *) "#   Energy (eV)    Transmission_coefficient"
        write(19, '(2G17.8)') (Transmission(i)%energy * hartree, Transmission(i)%tau, i = 1, size(Transmission))
        close(19)
        if (DUMP_FULL) then
            open(unit=19, file="TRANSMISSION_FULL")
            write(19, '(2ES25.17)') (Transmission(i)%energy * hartree, Transmission(i)%tau, i = 1, size(Transmission))
            close(19)
        end if
    end if
    if (if_LDOS) then
        write(16, *) "Writing LDOS..."
        open(unit=20, file="LDOS")
        write(20, '(4I5)') size(LDOS, 1), size(LDOS, 2), size(LDOS, 3), size(LDOS, 4)
        write(20, '(5G17.8)') ((((LDOS(i, j, k, ie), i = 1, size(LDOS, 1)), j = 1, size(LDOS, 2)), &
                               k = 1, size(LDOS, 3)), ie = 1, size(LDOS, 4))
        close(20)
        if (DUMP_FULL) then
            open(unit=20, file="LDOS_FULL")
            write(20, '(4I6)') size(LDOS, 1), size(LDOS, 2), size(LDOS, 3), size(LDOS, 4)
            write(20, '(ES25.17)') ((((LDOS(i, j, k, ie), i = 1, size(LDOS, 1)), j = 1, size(LDOS, 2)), &
                                    k = 1, size(LDOS, 3)), ie = 1, size(LDOS, 4))
            close(20)
        end if
    end if
    call negf_release()
    call wall(t_total1)
    write(16, *) "Wall time (Total): ", t_total1 - t_total0, "s"
    write(16, *) " |-- Real space to Plane wave (GPU): ", t_rtop, "s"
    write(16, *) " |-- Density (GPU inverse + accumulate): ", t_dens, "s"
    write(16, *) " |-- Density plane wave to real space (GPU): ", t_ptor, "s"
    write(16, *) " |-- Transmission (GPU): ", t_trans, "s"
    write(16, *) " |-- LDOS (GPU inverse): ", t_ldos, "s"
    write(16, *) " |-- LDOS plane wave to real space (GPU): ", t_ldos_ptor, "s"
    write(16, *) "********************"
    write(16, *) "*                  *"
    write(16, *) "*      DONE        *"
    write(16, *) "*                  *"
    write(16, *) "********************"
    close(16)

contains
    subroutine wall(t)
        real*8, intent(out) :: t
        integer(8) :: c, r
        call system_clock(c, r)
        t = dble(c) / dble(r)
    end subroutine wall

    subroutine banner(u)
        integer, intent(in) :: u
        character(len=8) :: d
        character(len=10) :: tm
        call date_and_time(d, tm)
        write(u, *) "**********************************************************************"
        write(u, *) " NEGF density / transmission / LDOS of a given potential"
        write(u, *) "   plane waves in x-y, finite differences in z; RGF energy loop on MI355X"
        write(u, *) " input files:  INPUT, POTENTIAL"
        write(u, *) " output files: OUTPUT, DENSITY, TRANSMISSION, LDOS"
        write(u, *) "**********************************************************************"
        write(u, *) "Calculated on ", d(1:4), "-", d(5:6), "-", d(7:8), " ", tm(1:2), ":", tm(3:4), ":", tm(5:6)
    end subroutine banner

    subroutine stop_run()
        write(16, *) "********************"
        write(16, *) "*                  *"
        write(16, *) "*      EXIT        *"
        write(16, *) "*                  *"
        write(16, *) "********************"
        close(16)
        stop 1
    end subroutine stop_run
end program negftrans_gpu
