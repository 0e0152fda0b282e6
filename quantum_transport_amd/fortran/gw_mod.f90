! gw_mod.f90 -- the RGF of the GW-SCBA loop on the MI355X.
!
! green_RGF_RS keeps the reference argument list (gw_mod.f90:1131-1139
! there): plain complex(8) (nm,nm,nx) arrays, real(8) cur, optional
! off-diagonal blocks.  The lead decimation uses z = E + i*1e-3, at most 100
! steps and the reset at nmax of gw_mod.f90:1400-1445 there; the right lead
! couples through H1i(:,:,nx).
module gw_mod
    use iso_c_binding
    use qt_capi
    implicit none
    private
    public :: green_RGF_RS

contains

    subroutine green_RGF_RS(TEMP, nm, nx, E, mu, Hii, H1i, sigma_lesser_ph, sigma_greater_ph, sigma_r_ph, &
                            ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1)
        integer, intent(in) :: nm, nx
        real(8), intent(in) :: E, mu(2), temp(2)
        complex(8), intent(in), dimension(nm, nm, nx), target :: sigma_lesser_ph, sigma_greater_ph, sigma_r_ph
        complex(8), intent(in), dimension(nm, nm, nx), target :: Hii, H1i
        complex(8), intent(inout), dimension(nm, nm, nx), target :: ldos, ndens, pdens
        real(8), intent(inout), dimension(nm, nm, nx), target :: cur
        complex(8), intent(inout), dimension(nm, nm, nx), optional, target :: GRi1, GLi1, GGi1
        real(8), intent(inout) :: tr, tre

        integer(c_int32_t), target :: nmv(nx), k1(1), info1(1)
        type(c_ptr), target :: hp(nx), cp(nx + 1), slp(nx), srp(nx)
        type(c_ptr), target :: pr(nx), pl(nx), pg(nx), pc(nx), p1(nx), p2(nx), p3(nx)
        real(c_double), target :: e1(1), tr1(1), tre1(1)
        type(c_ptr) :: ham, slptr, srptr
        type(qt_rgf_params) :: p
        type(qt_rgf_outputs) :: o
        integer(c_int) :: rc
        integer :: s

        call qt_ensure_init()
        do s = 1, nx
            nmv(s) = nm
            hp(s) = c_loc(Hii(1, 1, s))
            cp(s) = c_loc(H1i(1, 1, s))
            slp(s) = c_loc(sigma_lesser_ph(1, 1, s))
            srp(s) = c_loc(sigma_r_ph(1, 1, s))
            pr(s) = c_loc(ldos(1, 1, s))
            pl(s) = c_loc(ndens(1, 1, s))
            pg(s) = c_loc(pdens(1, 1, s))
            pc(s) = c_loc(cur(1, 1, s))
            if (present(GRi1)) p1(s) = c_loc(GRi1(1, 1, s))
            if (present(GLi1)) p2(s) = c_loc(GLi1(1, 1, s))
            if (present(GGi1)) p3(s) = c_loc(GGi1(1, 1, s))
        end do
        cp(nx + 1) = c_loc(H1i(1, 1, nx))
        slptr = c_loc(slp)
        srptr = c_loc(srp)
        if (all(sigma_lesser_ph == (0d0, 0d0)) .and. all(sigma_r_ph == (0d0, 0d0))) then
            slptr = c_null_ptr
            srptr = c_null_ptr
        end if
        p%variant = QT_RGF_GW
        p%flags = QT_F_LEAD_RESET
        p%lead_nmax = 100
        p%lead_eta = 1d-3
        p%mul = mu(1)
        p%mur = mu(2)
        p%templ = temp(1)
        p%tempr = temp(2)
        o%tr = c_loc(tr1)
        o%tre = c_loc(tre1)
        o%info = c_loc(info1)
        o%G_r = c_loc(pr)
        o%G_lesser = c_loc(pl)
        o%G_greater = c_loc(pg)
        o%cur = c_loc(pc)
        if (present(GRi1)) o%GRi1 = c_loc(p1)
        if (present(GLi1)) o%GLi1 = c_loc(p2)
        if (present(GGi1)) o%GGi1 = c_loc(p3)
        call qt_check(qt_ham_create(int(nx, c_int32_t), nmv, 1_c_int32_t, c_loc(hp), c_loc(cp), &
                                    c_null_ptr, ham), 'qt_ham_create')
        e1(1) = E
        k1(1) = 0
        rc = qt_rgf_solve(ham, p, 1_c_int32_t, e1, k1, slptr, srptr, o)
        call qt_check(qt_ham_destroy(ham), 'qt_ham_destroy')
        if (rc == QT_ERR_SINGULAR) then
            write (*, *) 'SEVERE warning: singular block inversion (block zeroed) at E =', E
        else
            call qt_check(rc, 'qt_rgf_solve')
        end if
        tr = tr1(1)
        tre = tre1(1)
    end subroutine green_RGF_RS

end module gw_mod
