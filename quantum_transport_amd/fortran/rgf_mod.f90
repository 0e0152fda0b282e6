! rgf_mod.f90 -- recursive Green's function solvers on the MI355X.
!
! Same public subroutines and argument lists as the reference rgf_mod
! (rgf_mod.f90:46-54, 272-280 there); the arithmetic runs in libqtrgf.so.
! A singular Schur-complement block is zeroed with a warning, as the
! reference's invert does (linalg.f90:50-58 there).
module rgf_mod
    use matrix_c, only: type_matrix_complex
    use qt_capi, only: QT_RGF_FORWARD, QT_RGF_BACKWARD
    use qt_rgf_host, only: rgf_point
    implicit none
    private
    integer, parameter :: dp = 8
    public :: rgf_variableblock_forward, rgf_variableblock_backward

contains

    subroutine rgf_variableblock_forward(nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sigma_lesser_ph, &
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
                       tr, tre, 'warn')
    end subroutine rgf_variableblock_forward

    subroutine rgf_variableblock_backward(nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sigma_lesser_ph, &
                                          sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, tr, tre)
        integer, intent(in) :: nx
        real(dp), intent(in) :: En, mul(:, :), mur(:, :), TEMPr(:, :), TEMPl(:, :)
        type(type_matrix_complex), intent(in), target :: Hii(nx), H1i(nx + 1), Sii(nx), &
                                                         sigma_lesser_ph(nx), sigma_r_ph(nx)
        type(type_matrix_complex), intent(inout), target :: G_greater(nx), G_lesser(nx), G_r(nx), &
                                                            Jdens(nx), Gl(nx), Gln(nx)
        real(dp), intent(out) :: tr, tre
        call rgf_point(QT_RGF_BACKWARD, 0, nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, &
                       sigma_lesser_ph, sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, &
                       tr, tre, 'warn')
    end subroutine rgf_variableblock_backward

end module rgf_mod
