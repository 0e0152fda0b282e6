! cuda_rgf_mod.f90 -- the GPU RGF entry points of the reference, on the
! MI355X library.
!
! Same public names and argument lists as the reference cuda_rgf_mod
! (cuda_rgf_mod.f90:32, 47, 53-59, 419-421 there).  The physics is that of
! rgf_mod's forward variant (the reference's CPU oracle); a singular block
! stops the program, as the reference GPU route does (:180-187 there).
! init/finish are reference counted: the reference calls them from every
! OpenMP thread of negf_solve.
module cuda_rgf_mod
    use iso_c_binding
    use matrix_c, only: type_matrix_complex, malloc, free
    use qt_capi, only: QT_RGF_FORWARD, qt_ensure_init, qt_finish, qt_check
    use qt_rgf_host, only: rgf_point
    implicit none
    private
    integer, parameter :: dp = 8
    integer, save :: users = 0
    public :: cuda_rgf_variableblock_forward
    public :: cuda_rgf_constblocksize, cuda_rgf_init, cuda_rgf_finish

contains

    subroutine cuda_rgf_init(nm)
        integer, intent(in) :: nm
        if (nm <= 0) error stop 'cuda_rgf_init: nm must be positive'
        !$omp critical (qt_cuda_rgf_users)
        call qt_ensure_init()
        users = users + 1
        !$omp end critical (qt_cuda_rgf_users)
    end subroutine cuda_rgf_init

    subroutine cuda_rgf_finish()
        !$omp critical (qt_cuda_rgf_users)
        users = max(users - 1, 0)
        !$omp end critical (qt_cuda_rgf_users)
    end subroutine cuda_rgf_finish

    subroutine cuda_rgf_constblocksize(nm, nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sigma_lesser_ph, &
                                       sigma_r_ph, G_r, G_lesser, G_greater, Jdens, tr, tre)
        integer, intent(in) :: nx, nm
        real(dp), intent(in) :: En, mul(:, :), mur(:, :), TEMPr(:, :), TEMPl(:, :)
        type(type_matrix_complex), intent(in), target :: Hii(nx), H1i(nx + 1), Sii(nx), &
                                                         sigma_lesser_ph(nx), sigma_r_ph(nx)
        type(type_matrix_complex), intent(inout), target :: G_greater(nx), G_lesser(nx), G_r(nx), Jdens(nx)
        real(dp), intent(out) :: tr, tre
        type(type_matrix_complex), allocatable, target :: Gl(:), Gln(:)
        integer :: s
        do s = 1, nx
            if (size(Hii(s)%m, 1) /= nm) error stop 'cuda_rgf_constblocksize: block size differs from nm'
        end do
        allocate (Gl(nx), Gln(nx))
        call rgf_point(QT_RGF_FORWARD, 0, nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, &
                       sigma_lesser_ph, sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, &
                       tr, tre, 'abort')
        call free(Gl)
        call free(Gln)
        deallocate (Gl, Gln)
    end subroutine cuda_rgf_constblocksize

    subroutine cuda_rgf_variableblock_forward(nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sigma_lesser_ph, &
                                              sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, tr, tre)
        integer, intent(in) :: nx
        real(dp), intent(in) :: En, mul(:, :), mur(:, :), TEMPr(:, :), TEMPl(:, :)
        type(type_matrix_complex), intent(in), target :: Hii(nx), H1i(nx + 1), Sii(nx), &
                                                         sigma_lesser_ph(nx), sigma_r_ph(nx)
        type(type_matrix_complex), intent(inout), target :: G_greater(nx), G_lesser(nx), G_r(nx), &
                                                            Jdens(nx), Gl(nx), Gln(nx)
        real(dp), intent(out) :: tr, tre
        call rgf_point(QT_RGF_FORWARD, 0, nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, &
                       sigma_lesser_ph, sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, &
                       tr, tre, 'abort')
    end subroutine cuda_rgf_variableblock_forward

end module cuda_rgf_mod
