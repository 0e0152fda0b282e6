! qt_rgf_host.f90 -- shared host logic of the Fortran RGF modules: turns
! type_matrix_complex block lists into the pointer tables of the C ABI and
! runs one batched device solve.
module qt_rgf_host
    use iso_c_binding
    use qt_capi
    use matrix_c, only: type_matrix_complex, ensure_block
    implicit none
    private
    public :: rgf_point, uniform_value, lists_identity, lists_zero

    real(8), parameter :: TOL_UNIFORM = 0d0

contains

    ! value of a matrix filled with one constant; stops otherwise (the
    ! device path takes one chemical potential / temperature per contact,
    ! as every caller fills mul/mur/TEMPl/TEMPr, negf_mod.f90:82-85 there)
    real(8) function uniform_value(A, name) result(v)
        real(8), intent(in) :: A(:, :)
        character(len=*), intent(in) :: name
        v = A(1, 1)
        if (any(abs(A - v) > TOL_UNIFORM)) then
            write (*, *) 'rgf: ', name, ' must be a uniform matrix on the device path'
            error stop 1
        end if
    end function uniform_value

    logical function lists_identity(list) result(b)
        type(type_matrix_complex), intent(in) :: list(:)
        integer :: s, i, j
        b = .true.
        do s = 1, size(list)
            do j = 1, size(list(s)%m, 2)
                do i = 1, size(list(s)%m, 1)
                    if (i == j) then
                        if (list(s)%m(i, j) /= (1d0, 0d0)) b = .false.
                    else
                        if (list(s)%m(i, j) /= (0d0, 0d0)) b = .false.
                    end if
                end do
            end do
            if (.not. b) return
        end do
    end function lists_identity

    logical function lists_zero(list) result(b)
        type(type_matrix_complex), intent(in) :: list(:)
        integer :: s
        b = .true.
        do s = 1, size(list)
            if (.not. allocated(list(s)%m)) cycle
            if (any(list(s)%m /= (0d0, 0d0))) then
                b = .false.
                return
            end if
        end do
    end function lists_zero

    ! One energy point through qt_rgf_solve with full block outputs.
    ! on_singular: 'warn' zeroes the block and warns (linalg.f90:50-58 there),
    ! 'abort' stops (cuda_rgf_mod.f90:180-187 there).
    subroutine rgf_point(variant, flags, nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, &
                         sigma_lesser_ph, sigma_r_ph, G_r, G_lesser, G_greater, Jdens, Gl, Gln, &
                         tr, tre, on_singular)
        integer(c_int32_t), intent(in) :: variant, flags
        integer, intent(in) :: nx
        real(8), intent(in) :: En, mul(:, :), mur(:, :), TEMPl(:, :), TEMPr(:, :)
        type(type_matrix_complex), intent(in), target :: Hii(nx), H1i(nx + 1), Sii(nx)
        type(type_matrix_complex), intent(in), target :: sigma_lesser_ph(nx), sigma_r_ph(nx)
        type(type_matrix_complex), intent(inout), target :: G_r(nx), G_lesser(nx), G_greater(nx)
        type(type_matrix_complex), intent(inout), target :: Jdens(nx), Gl(nx), Gln(nx)
        real(8), intent(out) :: tr, tre
        character(len=*), intent(in) :: on_singular

        integer(c_int32_t), target :: nm(nx), k1(1), info1(1), it1(2)
        type(c_ptr), target :: hp(nx), cp(nx + 1), sp(nx), slp(nx), srp(nx)
        type(c_ptr), target :: pr(nx), pl(nx), pg(nx), pj(nx), pgl(nx), pgn(nx)
        real(c_double), target :: e1(1), tr1(1), tre1(1)
        type(c_ptr) :: ham, sptr, slptr, srptr
        type(qt_rgf_params) :: p
        type(qt_rgf_outputs) :: o
        integer(c_int) :: rc
        integer :: s

        call qt_ensure_init()
        do s = 1, nx
            nm(s) = size(Hii(s)%m, 1)
            hp(s) = c_loc(Hii(s)%m)
            call ensure_block(G_r(s), nm(s), nm(s))
            call ensure_block(G_lesser(s), nm(s), nm(s))
            call ensure_block(G_greater(s), nm(s), nm(s))
            call ensure_block(Jdens(s), nm(s), nm(s))
            call ensure_block(Gl(s), nm(s), nm(s))
            call ensure_block(Gln(s), nm(s), nm(s))
            pr(s) = c_loc(G_r(s)%m)
            pl(s) = c_loc(G_lesser(s)%m)
            pg(s) = c_loc(G_greater(s)%m)
            pj(s) = c_loc(Jdens(s)%m)
            pgl(s) = c_loc(Gl(s)%m)
            pgn(s) = c_loc(Gln(s)%m)
        end do
        do s = 1, nx + 1
            cp(s) = c_loc(H1i(s)%m)
        end do
        sptr = c_null_ptr
        if (.not. lists_identity(Sii)) then
            do s = 1, nx
                sp(s) = c_loc(Sii(s)%m)
            end do
            sptr = c_loc(sp)
        end if
        slptr = c_null_ptr
        srptr = c_null_ptr
        if (.not. (lists_zero(sigma_lesser_ph) .and. lists_zero(sigma_r_ph))) then
            do s = 1, nx
                slp(s) = c_loc(sigma_lesser_ph(s)%m)
                srp(s) = c_loc(sigma_r_ph(s)%m)
            end do
            slptr = c_loc(slp)
            srptr = c_loc(srp)
        end if

        p%variant = variant
        p%flags = flags
        p%mul = uniform_value(mul, 'mul')
        p%mur = uniform_value(mur, 'mur')
        p%templ = uniform_value(TEMPl, 'TEMPl')
        p%tempr = uniform_value(TEMPr, 'TEMPr')
        o%tr = c_loc(tr1)
        o%tre = c_loc(tre1)
        o%info = c_loc(info1)
        o%lead_iters = c_loc(it1)
        o%G_r = c_loc(pr)
        o%G_lesser = c_loc(pl)
        o%G_greater = c_loc(pg)
        o%Jdens = c_loc(pj)
        o%Gl = c_loc(pgl)
        o%Gln = c_loc(pgn)

        call qt_check(qt_ham_create(int(nx, c_int32_t), nm, 1_c_int32_t, c_loc(hp), c_loc(cp), &
                                    sptr, ham), 'qt_ham_create')
        e1(1) = En
        k1(1) = 0
        rc = qt_rgf_solve(ham, p, 1_c_int32_t, e1, k1, slptr, srptr, o)
        call qt_check(qt_ham_destroy(ham), 'qt_ham_destroy')
        if (rc == QT_ERR_SINGULAR) then
            if (on_singular == 'abort') then
                write (*, *) 'rgf: singular block inversion at E =', En
                error stop 1
            end if
            write (*, *) 'SEVERE warning: singular block inversion (block zeroed) at E =', En
        else
            call qt_check(rc, 'qt_rgf_solve')
        end if
        if (iand(info1(1), QT_INFO_LEAD_NMAX) /= 0) &
            write (*, *) 'SEVERE warning: nmax reached in sancho!!! E =', En
        tr = tr1(1)
        tre = tre1(1)
    end subroutine rgf_point

end module qt_rgf_host
