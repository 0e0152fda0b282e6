! deviceHam_mod.f90 -- device Hamiltonian blocks from a Wannier ham_dat.
!
! devH_build_fromWannierFile keeps the reference argument list
! (deviceHam_mod.f90:234 there).  The file records are read in the
! reference order (header, cell, R-space elements "x y z m n Re Im",
! orbital centres); every slab gets the same H00 / H10 at kx = 0 and
! (ky, kz) = k(:, ik); Sii is the identity.  Block (i,k) of H00 sums
! Hr(x = i-k) e^{-i r.k}, block (i,k) of H10 sums Hr(x = i-k+ns) with the
! same phase r = (i-k) alpha + j beta + l gamma.  The reference's
! band-edge scan of the loader only prints CBM/VBM and is not repeated.
module deviceHam_mod
    use matrix_c, only: type_matrix_complex, malloc
    implicit none
    private
    integer, parameter :: dp = 8
    public :: devH_build_fromWannierFile

contains

    function cross3(a, b) result(c)
        real(dp), intent(in) :: a(3), b(3)
        real(dp) :: c(3)
        c = [a(2)*b(3) - a(3)*b(2), a(3)*b(1) - a(1)*b(3), a(1)*b(2) - a(2)*b(1)]
    end function cross3

    subroutine devH_build_fromWannierFile(fname, Hii, H1i, Sii, nx, nslab, nband, nk, k, length_x, &
                                          lreorder_axis, axis)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: nx, nslab, nk
        integer, intent(out) :: nband
        type(type_matrix_complex), dimension(:, :), intent(inout), allocatable :: Hii, H1i, Sii
        real(dp), intent(in) :: k(2, nk)
        real(dp), intent(out) :: length_x
        logical, intent(in), optional :: lreorder_axis
        integer, intent(in), optional :: axis(3)

        integer :: nvb, spin, n, i, r, nb, xmin, xmax, ymin, ymax, zmin, zmax, ik, s, kk, j, l, m, x0
        character(len=256) :: line
        real(dp) :: cell(3, 3), row(7), xhat(3), yhat(3), zhat(3), kv(3), rv(3)
        real(dp), allocatable :: ham(:, :)
        complex(dp), allocatable :: Hr(:, :, :, :, :), H00(:, :), H10(:, :)
        complex(dp) :: ph
        logical :: reorder

        reorder = .false.
        if (present(lreorder_axis)) reorder = lreorder_axis
        open (unit=10, file=trim(fname), status='old', action='read')
        read (10, *) nvb, spin
        read (10, '(A)') line
        do i = 1, 3
            read (10, *) cell(:, i)
        end do
        if (reorder) cell = cell(:, axis)
        read (10, '(A)') line
        read (10, *) n
        allocate (ham(n, 7))
        do r = 1, n
            read (10, *) row
            ham(r, :) = row
        end do
        if (reorder) ham(:, 1:3) = ham(:, axis)
        close (10)

        xmin = floor(minval(ham(:, 1))); xmax = floor(maxval(ham(:, 1)))
        ymin = floor(minval(ham(:, 2))); ymax = floor(maxval(ham(:, 2)))
        zmin = floor(minval(ham(:, 3))); zmax = floor(maxval(ham(:, 3)))
        nb = floor(maxval(ham(:, 4)))
        allocate (Hr(nb, nb, xmin:xmax, ymin:ymax, zmin:zmax))
        Hr = (0d0, 0d0)
        do r = 1, n
            Hr(floor(ham(r, 4)), floor(ham(r, 5)), floor(ham(r, 1)), floor(ham(r, 2)), floor(ham(r, 3))) = &
                cmplx(ham(r, 6), ham(r, 7), kind=dp)
        end do
        xhat = cell(:, 1)/norm2(cell(:, 1))
        yhat = -cross3(xhat, cell(:, 3))
        yhat = yhat/norm2(yhat)
        zhat = cell(:, 3)/norm2(cell(:, 3))
        length_x = abs(dot_product(cell(:, 1), xhat))
        nband = nb

        m = nb*nslab
        allocate (H00(m, m), H10(m, m))
        allocate (Hii(nx, nk), H1i(nx + 1, nk), Sii(nx, nk))
        do ik = 1, nk
            kv = 0d0*xhat + k(1, ik)*yhat + k(2, ik)*zhat
            H00 = (0d0, 0d0)
            H10 = (0d0, 0d0)
            do i = 1, nslab
                do kk = 1, nslab
                    do j = ymin, ymax
                        do l = zmin, zmax
                            rv = dble(i - kk)*cell(:, 1) + dble(j)*cell(:, 2) + dble(l)*cell(:, 3)
                            ph = exp(cmplx(0d0, -dot_product(rv, kv), kind=dp))
                            x0 = i - kk
                            if (x0 >= xmin .and. x0 <= xmax) &
                                H00((i - 1)*nb + 1:i*nb, (kk - 1)*nb + 1:kk*nb) = &
                                H00((i - 1)*nb + 1:i*nb, (kk - 1)*nb + 1:kk*nb) + Hr(:, :, x0, j, l)*ph
                            x0 = i - kk + nslab
                            if (x0 >= xmin .and. x0 <= xmax) &
                                H10((i - 1)*nb + 1:i*nb, (kk - 1)*nb + 1:kk*nb) = &
                                H10((i - 1)*nb + 1:i*nb, (kk - 1)*nb + 1:kk*nb) + Hr(:, :, x0, j, l)*ph
                        end do
                    end do
                end do
            end do
            do s = 1, nx
                call malloc(Hii(s, ik), m)
                call malloc(H1i(s, ik), m)
                call malloc(Sii(s, ik), m)
                Hii(s, ik)%m = H00
                H1i(s, ik)%m = H10
                do i = 1, m
                    Sii(s, ik)%m(i, i) = (1d0, 0d0)
                end do
            end do
            call malloc(H1i(nx + 1, ik), m)
            H1i(nx + 1, ik)%m = H10
        end do
    end subroutine devH_build_fromWannierFile

end module deviceHam_mod
