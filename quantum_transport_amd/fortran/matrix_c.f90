! matrix_c.f90 -- the block container of the RGF interfaces.
!
! type_matrix_complex is the allocatable column-major complex(8) block of
! the reference interfaces (matrix_c.f90:34-39 there), with the list
! allocation / release / size helpers its callers use.
module matrix_c
    implicit none
    private

    type, public :: type_matrix_complex
        complex(8), allocatable :: m(:, :)
        integer :: size(2) = 0
    end type type_matrix_complex

    public :: malloc, free, sizeof, trace, ensure_block

    interface malloc
        module procedure alloc_one, alloc_list_square, alloc_list_shapes
    end interface malloc

    interface free
        module procedure free_one
    end interface free

    interface sizeof
        module procedure list_sizes
    end interface sizeof

    interface trace
        module procedure block_trace, array_trace
    end interface trace

contains

    ! one block of n x nn (square when nn is absent), zero-filled
    pure subroutine alloc_one(M, n, nn)
        type(type_matrix_complex), intent(inout) :: M
        integer, intent(in) :: n
        integer, intent(in), optional :: nn
        integer :: c
        c = n
        if (present(nn)) c = nn
        call ensure_block(M, n, c)
        M%m = (0d0, 0d0)
    end subroutine alloc_one

    ! list of square blocks nm(i) x nm(i)
    pure subroutine alloc_list_square(list, nx, nm)
        integer, intent(in) :: nx, nm(nx)
        type(type_matrix_complex), intent(inout) :: list(nx)
        integer :: i
        do i = 1, nx
            call alloc_one(list(i), nm(i))
        end do
    end subroutine alloc_list_square

    ! list of blocks shaped nm(1,i) x nm(2,i)
    pure subroutine alloc_list_shapes(list, nx, nm)
        integer, intent(in) :: nx, nm(2, nx)
        type(type_matrix_complex), intent(inout) :: list(nx)
        integer :: i
        do i = 1, nx
            call alloc_one(list(i), nm(1, i), nm(2, i))
        end do
    end subroutine alloc_list_shapes

    ! (re)allocates M%m to r x c when its shape differs; keeps contents otherwise
    pure subroutine ensure_block(M, r, c)
        type(type_matrix_complex), intent(inout) :: M
        integer, intent(in) :: r, c
        if (allocated(M%m)) then
            if (size(M%m, 1) == r .and. size(M%m, 2) == c) then
                M%size = [r, c]
                return
            end if
            deallocate (M%m)
        end if
        allocate (M%m(r, c))
        M%m = (0d0, 0d0)
        M%size = [r, c]
    end subroutine ensure_block

    elemental subroutine free_one(M)
        type(type_matrix_complex), intent(inout) :: M
        if (allocated(M%m)) deallocate (M%m)
        M%size = 0
    end subroutine free_one

    ! shapes of a block list: nm(1:2, i)
    pure function list_sizes(list) result(nm)
        type(type_matrix_complex), intent(in) :: list(:)
        integer :: nm(2, size(list))
        integer :: i
        do i = 1, size(list)
            if (allocated(list(i)%m)) then
                nm(:, i) = shape(list(i)%m)
            else
                nm(:, i) = 0
            end if
        end do
    end function list_sizes

    pure function block_trace(M) result(t)
        type(type_matrix_complex), intent(in) :: M
        complex(8) :: t
        t = array_trace(M%m)
    end function block_trace

    pure function array_trace(A) result(t)
        complex(8), intent(in) :: A(:, :)
        complex(8) :: t
        integer :: i
        t = (0d0, 0d0)
        do i = 1, min(size(A, 1), size(A, 2))
            t = t + A(i, i)
        end do
    end function array_trace

end module matrix_c
