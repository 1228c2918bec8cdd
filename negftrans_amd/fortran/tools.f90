! tools.f90 -- LDOS output-grid helpers with the interface of src/tools.f90:6-40.
module tools
    implicit none
    private
    public LDOS_size, LDOS_grid_converter
contains
    ! Output extent along one axis: every point for selector 0, else a single point.
    integer function LDOS_size(LDOS_GRID_value, N_i)
        integer, intent(in) :: LDOS_GRID_value, N_i
        LDOS_size = merge(N_i, 1, LDOS_GRID_value == 0)
    end function LDOS_size

    ! Map grid index i to its output index ii and fold the axis rule into rescale:
    ! keep (0), integrate with weight delta_i (-1), select point i (= selector), drop others.
    subroutine LDOS_grid_converter(i, LDOS_GRID_value, delta_i, ii, rescale_factor)
        integer, intent(in) :: i, LDOS_GRID_value
        real*8, intent(in) :: delta_i
        integer, intent(out) :: ii
        real*8, intent(inout) :: rescale_factor
        ii = merge(i, 1, LDOS_GRID_value == 0)
        select case (LDOS_GRID_value)
        case (0)
        case (-1)
            rescale_factor = rescale_factor * delta_i
        case default
            if (LDOS_GRID_value /= i) rescale_factor = 0.D0
        end select
    end subroutine LDOS_grid_converter
end module tools
