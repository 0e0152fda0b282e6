! qt_capi.f90 -- iso_c_binding interface of libqtrgf.so (include/qt_rgf.h).
!
! The derived types mirror the C structs qt_rgf_params and qt_rgf_outputs
! field by field; every function returns the qt status code (0 = success).
module qt_capi
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: QT_OK = 0, QT_ERR_SINGULAR = 5
    integer(c_int32_t), parameter :: QT_RGF_FORWARD = 0, QT_RGF_BACKWARD = 1, QT_RGF_GW = 2
    integer(c_int32_t), parameter :: QT_PREC_FP64 = 0, QT_PREC_MIXED = 1, QT_PREC_MIXED47 = 2, &
                                     QT_PREC_FP64_3M = 3
    integer(c_int32_t), parameter :: QT_F_COUPLING_FIX = 1, QT_F_DEVICE_OUTPUTS = 2, QT_F_LEAD_RESET = 4
    integer(c_int32_t), parameter :: QT_INFO_SINGULAR = 1, QT_INFO_LEAD_NMAX = 2

    type, bind(C) :: qt_rgf_params
        integer(c_int32_t) :: variant = QT_RGF_FORWARD
        integer(c_int32_t) :: precision = QT_PREC_FP64
        integer(c_int32_t) :: flags = 0
        integer(c_int32_t) :: lead_nmax = 50
        real(c_double) :: mul = 0d0, mur = 0d0, templ = 300d0, tempr = 300d0
        real(c_double) :: lead_eta = 1d-5
        real(c_double) :: lead_tol = 1d-100
    end type qt_rgf_params

    type, bind(C) :: qt_rgf_outputs
        type(c_ptr) :: tr = c_null_ptr, tre = c_null_ptr, lead_iters = c_null_ptr, info = c_null_ptr
        type(c_ptr) :: gr_diag = c_null_ptr, gl_diag = c_null_ptr, gg_diag = c_null_ptr
        type(c_ptr) :: G_r = c_null_ptr, G_lesser = c_null_ptr, G_greater = c_null_ptr
        type(c_ptr) :: Jdens = c_null_ptr, Gl = c_null_ptr, Gln = c_null_ptr
        type(c_ptr) :: cur = c_null_ptr, GRi1 = c_null_ptr, GLi1 = c_null_ptr, GGi1 = c_null_ptr
    end type qt_rgf_outputs

    interface
        integer(c_int) function qt_init(device) bind(C, name="qt_init")
            import :: c_int
            integer(c_int), value :: device
        end function qt_init

        integer(c_int) function qt_finish() bind(C, name="qt_finish")
            import :: c_int
        end function qt_finish

        type(c_ptr) function qt_status_string_c(status) bind(C, name="qt_status_string")
            import :: c_int, c_ptr
            integer(c_int), value :: status
        end function qt_status_string_c

        integer(c_int) function qt_ham_create(nx, nm, nk, Hii, H1i, Sii, out) bind(C, name="qt_ham_create")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: nx, nk
            integer(c_int32_t), intent(in) :: nm(*)
            type(c_ptr), value :: Hii, H1i, Sii
            type(c_ptr), intent(out) :: out
        end function qt_ham_create

        integer(c_int) function qt_ham_destroy(h) bind(C, name="qt_ham_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function qt_ham_destroy

        integer(c_int) function qt_rgf_solve(h, p, npts, En, kidx, sigma_lesser_ph, sigma_r_ph, out) &
            bind(C, name="qt_rgf_solve")
            import :: c_int, c_int32_t, c_ptr, c_double, qt_rgf_params, qt_rgf_outputs
            type(c_ptr), value :: h
            type(qt_rgf_params), intent(in) :: p
            integer(c_int32_t), value :: npts
            real(c_double), intent(in) :: En(*)
            integer(c_int32_t), intent(in) :: kidx(*)
            type(c_ptr), value :: sigma_lesser_ph, sigma_r_ph
            type(qt_rgf_outputs), intent(inout) :: out
        end function qt_rgf_solve

        integer(c_int) function qt_sancho(nm, npts, En, eta, nmax, tol, reset_on_fail, S00, H00, H10, &
                                          G00, GBB, iters, converged) bind(C, name="qt_sancho")
            import :: c_int, c_int32_t, c_ptr, c_double
            integer(c_int32_t), value :: nm, npts, nmax, reset_on_fail
            real(c_double), intent(in) :: En(*)
            real(c_double), value :: eta, tol
            type(c_ptr), value :: S00, H00, H10, G00, GBB, iters, converged
        end function qt_sancho
    end interface

contains

    ! status text of a qt status code
    function qt_status_string(status) result(s)
        integer(c_int), intent(in) :: status
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: cs(:)
        type(c_ptr) :: p
        integer :: n
        p = qt_status_string_c(status)
        call c_f_pointer(p, cs, [256])
        n = 0
        do while (n < 256)
            if (cs(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate (character(len=n) :: s)
        s = transfer(cs(1:n), s)
    end function qt_status_string

    ! binds the process to its device on first use (QT_DEVICE, default 0)
    subroutine qt_ensure_init()
        character(len=16) :: v
        integer :: dev, ln, st
        integer(c_int) :: rc
        logical, save :: done = .false.
        if (done) return
        dev = 0
        call get_environment_variable("QT_DEVICE", v, ln, st)
        if (st == 0 .and. ln > 0) read (v, *) dev
        rc = qt_init(int(dev, c_int))
        if (rc /= QT_OK) then
            write (*, *) 'qt_init failed: ', qt_status_string(rc)
            error stop 1
        end if
        done = .true.
    end subroutine qt_ensure_init

    subroutine qt_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= QT_OK) then
            write (*, *) trim(where), ': ', qt_status_string(rc)
            error stop 1
        end if
    end subroutine qt_check

end module qt_capi
