! selfenergy_mod.f90 -- contact self-energies on the MI355X.
!
! sancho has the argument list of the reference's (private) Sancho-Rubio
! routine (rgf_mod.f90:482 there): surface and bulk Green's functions of a
! semi-infinite lead, z = E + i*1e-5, at most 50 decimation steps.
! sancho_batch solves many energies in one device call.
module selfenergy_mod
    use iso_c_binding
    use qt_capi
    implicit none
    private
    integer, parameter :: dp = 8
    public :: sancho, sancho_batch

contains

    subroutine sancho(nm, E, S00, H00, H10, G00, GBB)
        integer, intent(in) :: nm
        real(dp), intent(in) :: E
        complex(dp), intent(in), target :: S00(nm, nm), H00(nm, nm), H10(nm, nm)
        complex(dp), intent(out), target :: G00(nm, nm), GBB(nm, nm)
        real(dp) :: En(1)
        complex(dp), allocatable :: G3(:, :, :), B3(:, :, :)
        allocate (G3(nm, nm, 1), B3(nm, nm, 1))
        En(1) = E
        call sancho_batch(nm, 1, En, S00, reshape(H00, [nm, nm, 1]), reshape(H10, [nm, nm, 1]), &
                          G3, B3, 1d-5, 50)
        G00 = G3(:, :, 1)
        GBB = B3(:, :, 1)
    end subroutine sancho

    ! H00s(nm,nm,npts), H10s(nm,nm,npts) -> G00s, GBBs (nm,nm,npts)
    subroutine sancho_batch(nm, npts, En, S00, H00s, H10s, G00s, GBBs, eta, nmax, iters)
        integer, intent(in) :: nm, npts, nmax
        real(dp), intent(in) :: En(npts), eta
        complex(dp), intent(in), target :: S00(nm, nm), H00s(nm, nm, npts), H10s(nm, nm, npts)
        complex(dp), intent(out), target :: G00s(nm, nm, npts), GBBs(nm, nm, npts)
        integer, intent(out), optional :: iters(npts)
        type(c_ptr), target :: ph(npts), pc(npts), pg(npts), pb(npts)
        integer(c_int32_t), target :: it(npts), cv(npts)
        logical :: ident
        integer :: i, j, ip
        call qt_ensure_init()
        do ip = 1, npts
            ph(ip) = c_loc(H00s(1, 1, ip))
            pc(ip) = c_loc(H10s(1, 1, ip))
            pg(ip) = c_loc(G00s(1, 1, ip))
            pb(ip) = c_loc(GBBs(1, 1, ip))
        end do
        ident = .true.
        do j = 1, nm
            do i = 1, nm
                if (i == j .and. S00(i, j) /= (1d0, 0d0)) ident = .false.
                if (i /= j .and. S00(i, j) /= (0d0, 0d0)) ident = .false.
            end do
        end do
        if (ident) then
            call qt_check(qt_sancho(int(nm, c_int32_t), int(npts, c_int32_t), En, eta, int(nmax, c_int32_t), &
                                    1d-100, 0_c_int32_t, c_null_ptr, c_loc(ph), c_loc(pc), c_loc(pg), &
                                    c_loc(pb), c_loc(it), c_loc(cv)), 'qt_sancho')
        else
            call qt_check(qt_sancho(int(nm, c_int32_t), int(npts, c_int32_t), En, eta, int(nmax, c_int32_t), &
                                    1d-100, 0_c_int32_t, c_loc(S00), c_loc(ph), c_loc(pc), c_loc(pg), &
                                    c_loc(pb), c_loc(it), c_loc(cv)), 'qt_sancho')
        end if
        do ip = 1, npts
            if (cv(ip) == 0) write (*, *) 'SEVERE warning: nmax reached in sancho!!! E =', En(ip)
        end do
        if (present(iters)) iters = it
    end subroutine sancho_batch

end module selfenergy_mod
