! global.f90 -- derived types and physical constants of the NEGFtrans interface
! (the same type names, components and constant values as src/global.f90:5-31, so
! callers written against the reference compile unchanged).
module global
    implicit none

    ! all components in atomic units after the INPUT conversion (src/main.f90:182-198)
    type :: t_parameters
        real*8 :: V_L, V_R, MU, ETA, TEMPERATURE, LX, LY, LZ, ENCUT, GAP, VDS
        real*8 :: TRANSMISSION_GRID(2), LDOS_ENERGY_GRID(2)
    end type
    type :: t_timer
        real :: start, end, sum = 0.0
    end type
    type :: t_kpointmesh
        real*8 :: kx, ky, weight
    end type
    type :: t_transmission
        real*8 :: energy, tau
    end type
    type :: t_gfunc
        complex*16 :: diagonal, first_column, last_column
    end type

    ! CODATA values used by the reference (eV, Angstrom, Hartree/K)
    real*8, parameter :: hartree = 27.211386245988D0
    real*8, parameter :: k_B = 8.617333262D-5 / hartree
    real*8, parameter :: a_0 = 0.529177210903D0
    real*8, parameter :: pi = 3.141592653589793D0

    logical, parameter :: DEBUG = .false.
end module global
