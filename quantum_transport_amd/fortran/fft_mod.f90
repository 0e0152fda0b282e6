! fft_mod.f90 -- energy-axis correlation and convolution on the device.
!
! Keeps the interface of the reference's fft_mod (fft_mod.f90:58-168):
! corr1d2(n, X, Y, method) and conv1d2(n, X, Y, method) return Z(n).  Only
! method 'fft' is provided (the one gw_mod.f90:464 selects); its result is
! the zero-padded linear correlation / convolution, which the library computes
! with an LDS-resident FFT per series.  The batched forms take X(n, nseries)
! with the energy axis fastest, the layout of G_lesser_by_blocks(:, ik, ix, i)
! (gw_mod.f90:246-255), and process all series in one call.
module fft_mod
    use iso_c_binding
    use qt_capi
    implicit none
    private

    public :: corr1d2, conv1d2, corr1d2_batched, conv1d2_batched

    interface
        integer(c_int) function qt_corr1d2_batched(n, nseries, X, Y, Z, flags) &
            bind(C, name="qt_corr1d2_batched")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            integer(c_int32_t), value :: n, flags
            integer(c_int64_t), value :: nseries
            type(c_ptr), value :: X, Y, Z
        end function qt_corr1d2_batched

        integer(c_int) function qt_conv1d2_batched(n, nseries, X, Y, Z, flags) &
            bind(C, name="qt_conv1d2_batched")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            integer(c_int32_t), value :: n, flags
            integer(c_int64_t), value :: nseries
            type(c_ptr), value :: X, Y, Z
        end function qt_conv1d2_batched
    end interface

contains

    subroutine require_fft(method)
        character(len=*), intent(in) :: method
        if (trim(method) /= 'fft') then
            write (*, *) 'fft_mod: method "', trim(method), '" is not provided (use ''fft'')'
            error stop 1
        end if
    end subroutine require_fft

    ! Z(i) = sum_e X(e) Y(e - l), l = i - 1 - n/2  (fft_mod.f90:58-91)
    function corr1d2(n, X, Y, method) result(Z)
        integer, intent(in) :: n
        character(len=*), intent(in) :: method
        complex(8), intent(in), target :: X(n), Y(n)
        complex(8), target :: Z(n)
        call require_fft(method)
        call qt_ensure_init()
        call qt_check(qt_corr1d2_batched(int(n, c_int32_t), 1_c_int64_t, c_loc(X), c_loc(Y), &
                                         c_loc(Z), 0_c_int32_t), 'corr1d2')
    end function corr1d2

    ! Z(e) = sum_l X(e - l) Y(l + n/2 + 1)  (fft_mod.f90:129-168)
    function conv1d2(n, X, Y, method) result(Z)
        integer, intent(in) :: n
        character(len=*), intent(in) :: method
        complex(8), intent(in), target :: X(n), Y(n)
        complex(8), target :: Z(n)
        call require_fft(method)
        call qt_ensure_init()
        call qt_check(qt_conv1d2_batched(int(n, c_int32_t), 1_c_int64_t, c_loc(X), c_loc(Y), &
                                         c_loc(Z), 0_c_int32_t), 'conv1d2')
    end function conv1d2

    ! Z(:, s) = corr1d2(n, X(:, s), Y(:, s), 'fft') for every column s
    subroutine corr1d2_batched(n, nseries, X, Y, Z)
        integer, intent(in) :: n, nseries
        complex(8), intent(in), target :: X(n, nseries), Y(n, nseries)
        complex(8), intent(out), target :: Z(n, nseries)
        call qt_ensure_init()
        call qt_check(qt_corr1d2_batched(int(n, c_int32_t), int(nseries, c_int64_t), c_loc(X), &
                                         c_loc(Y), c_loc(Z), 0_c_int32_t), 'corr1d2_batched')
    end subroutine corr1d2_batched

    ! Z(:, s) = conv1d2(n, X(:, s), Y(:, s), 'fft') for every column s
    subroutine conv1d2_batched(n, nseries, X, Y, Z)
        integer, intent(in) :: n, nseries
        complex(8), intent(in), target :: X(n, nseries), Y(n, nseries)
        complex(8), intent(out), target :: Z(n, nseries)
        call qt_ensure_init()
        call qt_check(qt_conv1d2_batched(int(n, c_int32_t), int(nseries, c_int64_t), c_loc(X), &
                                         c_loc(Y), c_loc(Z), 0_c_int32_t), 'conv1d2_batched')
    end subroutine conv1d2_batched

end module fft_mod
