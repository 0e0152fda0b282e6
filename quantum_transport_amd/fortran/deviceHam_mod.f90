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
!
! devH_build_fromCOOfile keeps the reference argument list (deviceHam_mod.f90:
! 91 there) and its file format (load_COOmatrix, :40-89: "i j re [im]" per
! line, 0- or 1-based), for the fixed-blocksize route main.f90:113-114 takes.
! The block lists are (nx, 1), the shape negf_solve and main.f90 use (the
! reference's dummies are rank 1, which main.f90's rank-2 actuals do not
! match).  The blocks follow one Hermitian matrix given by its upper triangle
! (entries inside a diagonal block are mirrored, as the reference does;
! entries below the diagonal are accepted too).  Different from the
! reference, whose block does not run as written: row r (1-based) belongs to
! slab (r-1)/blocksize + 1 (r/blocksize + 1 there puts the last row of each
! slab into the next one at local row 0); H1i(s) holds H_{s,s-1} =
! H_{s-1,s}^H as the solver reads it (H_{s,s+1} there, with the last coupling
! left zero); H1i(1) = H1i(2) and H1i(nx+1) = H1i(nx) for the leads; the
! matrix order counts 0-based indices correctly.  The threshold filter and the
! graph slicing are commented out in the reference and not provided: ext and
! contactBlockSize are accepted and unused, blocksize is required.
module deviceHam_mod
    use matrix_c, only: type_matrix_complex, malloc
    implicit none
    private
    integer, parameter :: dp = 8
    public :: devH_build_fromWannierFile, devH_build_fromCOOfile, load_COOmatrix

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

    subroutine load_COOmatrix(fname, H, nnz, nm, row, col, use0index, iscomplex)
        character(len=*), intent(in) :: fname
        complex(dp), allocatable, intent(inout), dimension(:) :: H
        integer, allocatable, intent(inout), dimension(:) :: row, col
        integer, intent(out) :: nnz, nm
        logical, intent(in), optional :: use0index, iscomplex
        logical :: zero_based, cplx
        integer :: fu, io, i, j, k
        real(dp) :: re, im
        character(len=512) :: line
        zero_based = .false.
        if (present(use0index)) zero_based = use0index
        cplx = .false.
        if (present(iscomplex)) cplx = iscomplex
        open (newunit=fu, file=trim(fname), status='old', action='read')
        nnz = 0
        do
            read (fu, '(A)', iostat=io) line
            if (io /= 0) exit
            if (len_trim(line) > 0) nnz = nnz + 1
        end do
        if (allocated(H)) deallocate (H)
        if (allocated(row)) deallocate (row)
        if (allocated(col)) deallocate (col)
        allocate (H(nnz), row(nnz), col(nnz))
        rewind (fu)
        k = 0
        nm = 0
        do while (k < nnz)
            read (fu, '(A)') line
            if (len_trim(line) == 0) cycle
            k = k + 1
            im = 0d0
            if (cplx) then
                read (line, *) i, j, re, im
            else
                read (line, *) i, j, re
            end if
            if (zero_based) then
                i = i + 1
                j = j + 1
            end if
            row(k) = i
            col(k) = j
            H(k) = dcmplx(re, im)
            nm = max(nm, i, j)
        end do
        close (fu)
    end subroutine load_COOmatrix

    subroutine devH_build_fromCOOfile(hfname, Hii, H1i, Sii, ext, contactBlockSize, nx, use0index, &
                                      iscomplex, threshold, blocksize, sfname)
        character(len=*), intent(in) :: hfname
        character(len=*), intent(in), optional :: sfname
        integer, intent(in) :: ext(2), contactBlockSize(2)
        integer, intent(out) :: nx
        integer, intent(in), optional :: blocksize
        type(type_matrix_complex), dimension(:, :), intent(inout), allocatable :: Hii, H1i, Sii
        logical, intent(in), optional :: use0index, iscomplex
        real(dp), intent(in), optional :: threshold
        complex(dp), allocatable :: H(:), S(:)
        integer, allocatable :: row(:), col(:), srow(:), scol(:)
        integer :: nnz, norb, snnz, snorb, bs, k, si, sj, mi, mj
        if (.not. present(blocksize)) error stop 'devH_build_fromCOOfile: blocksize is required'
        bs = blocksize
        call load_COOmatrix(hfname, H, nnz, norb, row, col, use0index, iscomplex)
        if (bs <= 0 .or. mod(norb, bs) /= 0) error stop 'devH_build_fromCOOfile: order not a multiple of blocksize'
        nx = norb/bs
        if (nx < 2) error stop 'devH_build_fromCOOfile: fewer than two slabs'
        allocate (Hii(nx, 1), H1i(nx + 1, 1), Sii(nx, 1))
        do k = 1, nx
            call malloc(Hii(k, 1), bs)
            call malloc(Sii(k, 1), bs)
        end do
        do k = 1, nx + 1
            call malloc(H1i(k, 1), bs)
        end do
        do k = 1, nnz
            si = (row(k) - 1)/bs + 1
            sj = (col(k) - 1)/bs + 1
            mi = row(k) - (si - 1)*bs
            mj = col(k) - (sj - 1)*bs
            if (si == sj) then
                Hii(si, 1)%m(mi, mj) = H(k)
                Hii(si, 1)%m(mj, mi) = conjg(H(k))
            else if (sj == si + 1) then      ! H_{si,si+1}: H1i(sj) = its conjugate transpose
                H1i(sj, 1)%m(mj, mi) = conjg(H(k))
            else if (sj == si - 1) then      ! H_{si,si-1}
                H1i(si, 1)%m(mi, mj) = H(k)
            else
                error stop 'devH_build_fromCOOfile: the matrix is not block-tridiagonal at this blocksize'
            end if
        end do
        H1i(1, 1)%m = H1i(2, 1)%m
        H1i(nx + 1, 1)%m = H1i(nx, 1)%m
        if (present(sfname)) then
            call load_COOmatrix(sfname, S, snnz, snorb, srow, scol, use0index, iscomplex)
            if (snorb > norb) error stop 'devH_build_fromCOOfile: overlap larger than the Hamiltonian'
            do k = 1, snnz
                si = (srow(k) - 1)/bs + 1
                sj = (scol(k) - 1)/bs + 1
                if (si /= sj) error stop 'devH_build_fromCOOfile: the overlap must be block diagonal'
                mi = srow(k) - (si - 1)*bs
                mj = scol(k) - (sj - 1)*bs
                Sii(si, 1)%m(mi, mj) = S(k)
                Sii(si, 1)%m(mj, mi) = conjg(S(k))
            end do
        else
            do k = 1, nx
                do mi = 1, bs
                    Sii(k, 1)%m(mi, mi) = (1d0, 0d0)
                end do
            end do
        end if
    end subroutine devH_build_fromCOOfile

end module deviceHam_mod
