! gw_mod.f90 -- the RGF of the GW-SCBA loop on the MI355X.
!
! green_RGF_RS keeps the reference argument list (gw_mod.f90:1131-1139
! there): plain complex(8) (nm,nm,nx) arrays, real(8) cur, optional
! off-diagonal blocks.  The lead decimation uses z = E + i*1e-3, at most 100
! steps and the reset at nmax of gw_mod.f90:1400-1445 there; the right lead
! couples through H1i(:,:,nx).
!
! The SCBA stages of gw_ephoton_3d_ijs are provided as device calls on the
! reference's buffers: gw_polarization (P loop, gw_mod.f90:444-550),
! gw_selfenergy (Sigma loop, :641-743), gw_mix (:747-749) on the
! element-major by_blocks arrays (nen, nk, nx, nm*nm), and green_rgf_calc_w
! (the W solve called at :597) on the energy-major (nm, nm, nx, nw) blocks,
! and calc_sigma_ephoton_monochromatic_ext (:1075-1122) with its argument list.
! calc_block_current and calc_block_collision keep the argument lists of
! their call sites (:384, :795); the reference does not define them, the
! definitions are the library's (qt_gw_block_current, qt_gw_block_collision).
! gw_ephoton_3d_ijs (:78-1060) keeps its argument list and runs the SCBA
! loop with the self-energies, G and P on the device (qt_gw_scba_step) for
! one process, one k point and no e-photon coupling.
module gw_mod
    use iso_c_binding
    use qt_capi
    implicit none
    private
    public :: green_RGF_RS, green_rgf_calc_w, gw_polarization, gw_selfenergy, gw_mix, &
              calc_sigma_ephoton_monochromatic_ext, calc_block_current, calc_block_collision, gw_ephoton_3d_ijs

    interface
        integer(c_int) function qt_gw_polarization(nen, nky, nkz, nx, nm, ndiag, dEn, spindeg, Gr, Gl, Gg, &
                                                   Pr, Pl, Pg, flags) bind(C, name="qt_gw_polarization")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nen, nky, nkz, nx, nm, ndiag, flags
            real(c_double), value :: dEn, spindeg
            type(c_ptr), value :: Gr, Gl, Gg, Pr, Pl, Pg
        end function qt_gw_polarization

        integer(c_int) function qt_gw_selfenergy(nen, nky, nkz, nx, nm, ndiag, dEn, Gr, Gl, Gg, Wr, Wl, Wg, &
                                                 Sr, Sl, Sg, flags) bind(C, name="qt_gw_selfenergy")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nen, nky, nkz, nx, nm, ndiag, flags
            real(c_double), value :: dEn
            type(c_ptr), value :: Gr, Gl, Gg, Wr, Wl, Wg, Sr, Sl, Sg
        end function qt_gw_selfenergy

        integer(c_int) function qt_gw_mix(n, alpha, S, Snew, flags) bind(C, name="qt_gw_mix")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            integer(c_int64_t), value :: n
            real(c_double), value :: alpha
            type(c_ptr), value :: S, Snew
            integer(c_int32_t), value :: flags
        end function qt_gw_mix

        integer(c_int) function qt_gw_wsolve_batched(nm, nx, nw, Vii, V1i, PR, PL, PG, WR, WL, WG, flags) &
            bind(C, name="qt_gw_wsolve_batched")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: nm, nx, nw, flags
            type(c_ptr), value :: Vii, V1i, PR, PL, PG, WR, WL, WG
        end function qt_gw_wsolve_batched

        integer(c_int) function qt_gw_sigma_ephoton(nm, length, nen, nop, pre_fact, intensity, hw, Mii, &
                                                    Gl, Gg, Sr, Sl, Sg, flags) bind(C, name="qt_gw_sigma_ephoton")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nm, length, nen, nop, flags
            real(c_double), value :: pre_fact, intensity, hw
            type(c_ptr), value :: Mii, Gl, Gg, Sr, Sl, Sg
        end function qt_gw_sigma_ephoton

        integer(c_int) function qt_gw_block_current(nen, nx, nb, ns, En, spindeg, Hii, GL, cur, jdens, &
                                                    tot_cur, tot_ecur) bind(C, name="qt_gw_block_current")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nen, nx, nb, ns
            type(c_ptr), value :: En
            real(c_double), value :: spindeg
            type(c_ptr), value :: Hii, GL, cur, jdens, tot_cur, tot_ecur
        end function qt_gw_block_current

        integer(c_int) function qt_gw_scba_step(h, prm, nen, En, ndiag, spindeg, alpha, Vii, V1i, S, G, Pol, &
                                                tr, tre, cur) bind(C, name="qt_gw_scba_step")
            import :: c_int, c_int32_t, c_double, c_ptr, qt_rgf_params
            type(c_ptr), value :: h
            type(qt_rgf_params), intent(in) :: prm
            integer(c_int32_t), value :: nen, ndiag
            type(c_ptr), value :: En
            real(c_double), value :: spindeg, alpha
            type(c_ptr), value :: Vii, V1i, S, G, Pol, tr, tre, cur
        end function qt_gw_scba_step

        integer(c_int) function qt_device_alloc(bytes, out) bind(C, name="qt_device_alloc")
            import :: c_int, c_size_t, c_ptr
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: out
        end function qt_device_alloc

        integer(c_int) function qt_device_free(ptr) bind(C, name="qt_device_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
        end function qt_device_free

        integer(c_int) function qt_memcpy_h2d(dst, src, bytes) bind(C, name="qt_memcpy_h2d")
            import :: c_int, c_size_t, c_ptr
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function qt_memcpy_h2d

        integer(c_int) function qt_memcpy_d2h(dst, src, bytes) bind(C, name="qt_memcpy_d2h")
            import :: c_int, c_size_t, c_ptr
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function qt_memcpy_d2h

        integer(c_int) function qt_gw_block_collision(nen, nx, nm, En, spindeg, Sl, Sg, Gl, Gg, Itot, Ispec) &
            bind(C, name="qt_gw_block_collision")
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nen, nx, nm
            type(c_ptr), value :: En
            real(c_double), value :: spindeg
            type(c_ptr), value :: Sl, Sg, Gl, Gg, Itot, Ispec
        end function qt_gw_block_collision
    end interface

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

    ! P^<, P^>, P^R of every in-band element (|row - col| <= ndiag) from
    ! G^R, G^<, G^> (i*Im G is used), k-sums over nky*nkz points; dEn = En(2)-En(1)
    subroutine gw_polarization(nen, nky, nkz, nx, nm, ndiag, dEn, spindeg, G_r, G_lesser, G_greater, &
                               P_retarded, P_lesser, P_greater)
        integer, intent(in) :: nen, nky, nkz, nx, nm, ndiag
        real(8), intent(in) :: dEn, spindeg
        complex(8), intent(in), dimension(nen, nky*nkz, nx, nm*nm), target :: G_r, G_lesser, G_greater
        complex(8), intent(out), dimension(nen, nky*nkz, nx, nm*nm), target :: P_retarded, P_lesser, P_greater
        call qt_ensure_init()
        call qt_check(qt_gw_polarization(int(nen, c_int32_t), int(nky, c_int32_t), int(nkz, c_int32_t), &
                                         int(nx, c_int32_t), int(nm, c_int32_t), int(ndiag, c_int32_t), &
                                         real(dEn, c_double), real(spindeg, c_double), c_loc(G_r), &
                                         c_loc(G_lesser), c_loc(G_greater), c_loc(P_retarded), &
                                         c_loc(P_lesser), c_loc(P_greater), 0_c_int32_t), 'gw_polarization')
    end subroutine gw_polarization

    ! new Sigma^R, Sigma^<, Sigma^> (band zeroing and (S + S^T)/2 included)
    subroutine gw_selfenergy(nen, nky, nkz, nx, nm, ndiag, dEn, G_r, G_lesser, G_greater, W_retarded, &
                             W_lesser, W_greater, Sigma_r_new, Sigma_lesser_new, Sigma_greater_new)
        integer, intent(in) :: nen, nky, nkz, nx, nm, ndiag
        real(8), intent(in) :: dEn
        complex(8), intent(in), dimension(nen, nky*nkz, nx, nm*nm), target :: G_r, G_lesser, G_greater
        complex(8), intent(in), dimension(nen, nky*nkz, nx, nm*nm), target :: W_retarded, W_lesser, W_greater
        complex(8), intent(out), dimension(nen, nky*nkz, nx, nm*nm), target :: Sigma_r_new, Sigma_lesser_new, &
                                                                              Sigma_greater_new
        call qt_ensure_init()
        call qt_check(qt_gw_selfenergy(int(nen, c_int32_t), int(nky, c_int32_t), int(nkz, c_int32_t), &
                                       int(nx, c_int32_t), int(nm, c_int32_t), int(ndiag, c_int32_t), &
                                       real(dEn, c_double), c_loc(G_r), c_loc(G_lesser), c_loc(G_greater), &
                                       c_loc(W_retarded), c_loc(W_lesser), c_loc(W_greater), &
                                       c_loc(Sigma_r_new), c_loc(Sigma_lesser_new), c_loc(Sigma_greater_new), &
                                       0_c_int32_t), 'gw_selfenergy')
    end subroutine gw_selfenergy

    ! Sigma = Sigma + alpha_mix (Sigma_new - Sigma)
    subroutine gw_mix(n, alpha_mix, Sigma, Sigma_new)
        integer(8), intent(in) :: n
        real(8), intent(in) :: alpha_mix
        complex(8), intent(inout), target :: Sigma(n)
        complex(8), intent(in), target :: Sigma_new(n)
        call qt_ensure_init()
        call qt_check(qt_gw_mix(int(n, c_int64_t), real(alpha_mix, c_double), c_loc(Sigma), c_loc(Sigma_new), &
                                0_c_int32_t), 'gw_mix')
    end subroutine gw_mix

    ! e-photon self-energies in the monochromatic assumption:
    ! Sig^<> = pre_fact intensity / hw^2 M [G^<>(E - hw) + G^<>(E + hw)] M,
    ! Sig^R = i Im(Sig^> - Sig^<) / 2; G carries nop extra energies on each side
    subroutine calc_sigma_ephoton_monochromatic_ext(nm, length, nen, En, nop, Mii, G_lesser, G_greater, &
                                                    Sig_retarded, Sig_lesser, Sig_greater, pre_fact, &
                                                    intensity, hw)
        integer, intent(in) :: nm, length, nen, nop
        real(8), intent(in) :: En(nen), pre_fact, intensity, hw
        complex(8), intent(in), dimension(nm, nm, length), target :: Mii
        complex(8), intent(in), dimension(nm, nm, length, nen + 2 * nop), target :: G_lesser, G_greater
        complex(8), intent(inout), dimension(nm, nm, length, nen), target :: Sig_retarded, Sig_lesser, &
                                                                           Sig_greater
        call qt_ensure_init()
        call qt_check(qt_gw_sigma_ephoton(int(nm, c_int32_t), int(length, c_int32_t), int(nen, c_int32_t), &
                                          int(nop, c_int32_t), real(pre_fact, c_double), &
                                          real(intensity, c_double), real(hw, c_double), c_loc(Mii), &
                                          c_loc(G_lesser), c_loc(G_greater), c_loc(Sig_retarded), &
                                          c_loc(Sig_lesser), c_loc(Sig_greater), 0_c_int32_t), &
                      'calc_sigma_ephoton_monochromatic_ext')
    end subroutine calc_sigma_ephoton_monochromatic_ext

    ! Current matrix of every cell bond (cells of nb orbitals, ns per slab) and
    ! energy, and its energy integrals tot_cur, tot_ecur (A; the energy
    ! current weighted by E)
    subroutine calc_block_current(Hii, G_lesser, cur, nen, en, spindeg, nb, ns, nm, nx, tot_cur, tot_ecur, jdens)
        integer, intent(in) :: nen, nb, ns, nm, nx
        real(8), intent(in) :: en(nen), spindeg
        complex(8), intent(in), dimension(nm, nm, nx), target :: Hii
        complex(8), intent(in), dimension(nm, nm, nx, nen), target :: G_lesser
        real(8), intent(in), dimension(nm, nm, nx, nen), target :: cur
        real(8), intent(inout), dimension(nb, nb, nx * ns), target :: tot_cur, tot_ecur
        real(8), intent(inout), dimension(nb, nb, nx * ns, nen), target :: jdens
        real(8), target :: e(nen)
        if (nm /= nb * ns) then
            write (*, *) 'calc_block_current: nm must be nb*ns'
            error stop 1
        end if
        e = en
        call qt_ensure_init()
        call qt_check(qt_gw_block_current(int(nen, c_int32_t), int(nx, c_int32_t), int(nb, c_int32_t), &
                                          int(ns, c_int32_t), c_loc(e), real(spindeg, c_double), c_loc(Hii), &
                                          c_loc(G_lesser), c_loc(cur), c_loc(jdens), c_loc(tot_cur), &
                                          c_loc(tot_ecur)), 'calc_block_current')
    end subroutine calc_block_current

    ! Collision integrals Ispec = Sig^< G^> - Sig^> G^< per slab block and
    ! energy, Itot their energy integral (dE spindeg / (2 pi))
    subroutine calc_block_collision(Sig_lesser, Sig_greater, G_lesser, G_greater, nen, en, spindeg, nm, nx, &
                                    Itot, Ispec)
        integer, intent(in) :: nen, nm, nx
        real(8), intent(in) :: en(nen), spindeg
        complex(8), intent(in), dimension(nm, nm, nx, nen), target :: Sig_lesser, Sig_greater, G_lesser, &
                                                                    G_greater
        complex(8), intent(inout), dimension(nm, nm, nx), target :: Itot
        complex(8), intent(inout), dimension(nm, nm, nx, nen), target :: Ispec
        real(8), target :: e(nen)
        e = en
        call qt_ensure_init()
        call qt_check(qt_gw_block_collision(int(nen, c_int32_t), int(nx, c_int32_t), int(nm, c_int32_t), &
                                            c_loc(e), real(spindeg, c_double), c_loc(Sig_lesser), &
                                            c_loc(Sig_greater), c_loc(G_lesser), c_loc(G_greater), c_loc(Itot), &
                                            c_loc(Ispec)), 'calc_block_collision')
    end subroutine calc_block_collision

    ! The SCBA loop of gw_ephoton_3d_ijs (gw_mod.f90:78-1060 there) with its
    ! argument list, iterations iter = 0 .. niter: G by green_RGF_RS with the
    ! current self-energies, P, W, Sigma and mixing on the device
    ! (qt_gw_scba_step; Sigma, G and P stay there between iterations), then the
    ! reference's outputs of the iteration: gw_ldos, gw_Jdens, gw_trL, gw_trR,
    ! Id_iteration.dat, gw_I, gw_EI and gw_Scat (the collision integrals of the
    ! mixed self-energy).  Provided for one process (comm_size = 1,
    ! local_NE = nen), one k point (nky = nkz = 1), slabs of NB*NS = nm
    ! orbitals, and without the e-photon coupling (intensity = 0,
    ! labs = .false.); Pii, P1i, polarization and hw are then not read.
    subroutine gw_ephoton_3d_ijs(alpha_mix, niter, NB, NS, nm, nx, nky, nkz, ndiag, Lx, nen, en, temp, mu, &
                                 Hii, H1i, Vii, V1i, spindeg, Pii, P1i, polarization, intensity, hw, labs, &
                                 comm_size, comm_rank, local_NE, first_local_energy)
        use Output, only: write_spectrum_from_diagonals, write_transmission_spectrum_k, write_total_current, &
                          write_current_spectrum_summed_over_kz, write_current_summed_over_k, write_spectrum
        integer(kind=4), intent(in) :: comm_size, comm_rank, local_NE, first_local_energy
        integer, intent(in) :: nm, nx, nen, niter, NB, NS, ndiag, nky, nkz
        real(8), intent(in) :: en(nen), temp(2), mu(2), Lx, alpha_mix, spindeg
        complex(8), intent(in), dimension(nm, nm, nx, nky*nkz), target :: Hii, H1i, Vii, V1i
        real(8), intent(in) :: polarization(3), intensity, hw
        logical, intent(in) :: labs
        complex(8), intent(in) :: Pii(nm, nm, 3, nx, nky*nkz), P1i(nm, nm, 3, nx, nky*nkz)

        integer(c_int32_t), target :: nmv(nx)
        type(c_ptr), target :: hp(nx), cp(nx + 1), Sd(3), Gd(3), Pd(3)
        type(c_ptr) :: ham, dV(2)
        type(qt_rgf_params) :: p
        real(8), target :: e(nen), tr(1, nen), tre(1, nen)
        real(8), allocatable, target :: cur(:, :, :, :), jdens(:, :, :, :), tot_cur(:, :, :), tot_ecur(:, :, :)
        complex(8), allocatable, target :: stage(:, :, :), blk(:, :, :, :, :), D(:, :, :), Itot(:, :, :), &
                                           Ispec(:, :, :, :)
        integer(c_size_t) :: arr, vbytes
        integer(c_int) :: rc
        integer :: s, i, iter, ie, q

        if (comm_size /= 1 .or. local_NE /= nen .or. nky*nkz /= 1 .or. labs .or. intensity /= 0d0 .or. &
            nm /= NB*NS) then
            write (*, *) 'gw_ephoton_3d_ijs: provided for comm_size = 1, local_NE = nen, nky = nkz = 1, ', &
                'nm = NB*NS, intensity = 0 and labs = .false.'
            error stop 1
        end if
        call qt_ensure_init()
        e = en
        do s = 1, nx
            nmv(s) = nm
            hp(s) = c_loc(Hii(1, 1, s, 1))
            cp(s) = c_loc(H1i(1, 1, s, 1))
        end do
        cp(nx + 1) = c_loc(H1i(1, 1, nx, 1))  ! the right lead couples through H1i(nx)
        call qt_check(qt_ham_create(int(nx, c_int32_t), nmv, 1_c_int32_t, c_loc(hp), c_loc(cp), c_null_ptr, ham), &
                      'qt_ham_create')
        p%variant = QT_RGF_GW
        p%flags = QT_F_LEAD_RESET
        p%lead_nmax = 100
        p%lead_eta = 1d-3
        p%mul = mu(1)
        p%mur = mu(2)
        p%templ = temp(1)
        p%tempr = temp(2)
        ! device arrays: Sigma (zero at the start, gw_mod.f90:330-332 there), G, P,
        ! element-major: series (ij, ix) of nen energies, ij = row + col*nm, i.e.
        ! stage(ie, ix, ij) here; blk holds blocks by energy, (nm, nm, nx, nen)
        arr = int(16, c_size_t)*int(nm, c_size_t)*int(nm, c_size_t)*int(nx, c_size_t)*int(nen, c_size_t)
        vbytes = int(16, c_size_t)*int(nm, c_size_t)*int(nm, c_size_t)*int(nx, c_size_t)
        allocate (stage(nen, nx, nm*nm), blk(nm, nm, nx, nen, 4), D(nx*nm, 1, nen), cur(nm, nm, nx, nen))
        allocate (jdens(NB, NB, nx*NS, nen), tot_cur(NB, NB, nx*NS), tot_ecur(NB, NB, nx*NS))
        allocate (Itot(nm, nm, nx), Ispec(nm, nm, nx, nen))
        stage = (0d0, 0d0)
        do i = 1, 3
            call qt_check(qt_device_alloc(arr, Sd(i)), 'qt_device_alloc')
            call qt_check(qt_device_alloc(arr, Gd(i)), 'qt_device_alloc')
            call qt_check(qt_device_alloc(arr, Pd(i)), 'qt_device_alloc')
            call qt_check(qt_memcpy_h2d(Sd(i), c_loc(stage), arr), 'qt_memcpy_h2d')
        end do
        call qt_check(qt_device_alloc(vbytes, dV(1)), 'qt_device_alloc')
        call qt_check(qt_device_alloc(vbytes, dV(2)), 'qt_device_alloc')
        call qt_check(qt_memcpy_h2d(dV(1), c_loc(Vii), vbytes), 'qt_memcpy_h2d')
        call qt_check(qt_memcpy_h2d(dV(2), c_loc(V1i), vbytes), 'qt_memcpy_h2d')

        do iter = 0, niter
            rc = qt_gw_scba_step(ham, p, int(nen, c_int32_t), c_loc(e), int(ndiag, c_int32_t), &
                                 real(spindeg, c_double), real(alpha_mix, c_double), dV(1), dV(2), c_loc(Sd), &
                                 c_loc(Gd), c_loc(Pd), c_loc(tr), c_loc(tre), c_loc(cur))
            if (rc == QT_ERR_SINGULAR) then
                write (*, *) 'gw_ephoton_3d_ijs: singular block inversion (block zeroed), iter =', iter
            else
                call qt_check(rc, 'qt_gw_scba_step')
            end if
            ! G^R of this iteration: the diagonals of gw_ldos
            call qt_check(qt_memcpy_d2h(c_loc(stage), Gd(1), arr), 'qt_memcpy_d2h')
            do ie = 1, nen
                do s = 1, nx
                    do i = 1, nm
                        D((s - 1)*nm + i, 1, ie) = stage(ie, s, i + (i - 1)*nm)
                    end do
                end do
            end do
            call write_spectrum_from_diagonals('gw_ldos', iter, D, nen, en, 1, nx, NB, NS, Lx, (/1.0d0, -2.0d0/), &
                                               .false.)
            ! blocks by energy of G^<, G^> and of the mixed Sigma^<, Sigma^>
            do q = 1, 4
                if (q <= 2) then
                    call qt_check(qt_memcpy_d2h(c_loc(stage), Gd(q + 1), arr), 'qt_memcpy_d2h')
                else
                    call qt_check(qt_memcpy_d2h(c_loc(stage), Sd(q - 1), arr), 'qt_memcpy_d2h')
                end if
                do ie = 1, nen
                    do s = 1, nx
                        blk(:, :, s, ie, q) = reshape(stage(ie, s, :), (/nm, nm/))
                    end do
                end do
            end do
            call calc_block_current(Hii(:, :, :, 1), blk(:, :, :, :, 1), cur, nen, en, spindeg, NB, NS, nm, nx, &
                                    tot_cur, tot_ecur, jdens)
            call write_current_spectrum_summed_over_kz('gw_Jdens', iter, jdens, nen, en, nx*NS, NB, Lx, 1, .false.)
            call write_transmission_spectrum_k('gw_trL', iter, tr, nen, en, 1, .false.)
            call write_transmission_spectrum_k('gw_trR', iter, tre, nen, en, 1, .false.)
            call write_total_current('Id_iteration.dat', iter, sum(tr), sum(tre), en(2) - en(1), spindeg, 1)
            call write_current_summed_over_k('gw_I', iter, tot_cur, nx*NS, NB, Lx, 1)
            call write_current_summed_over_k('gw_EI', iter, tot_ecur, nx*NS, NB, Lx, 1)
            call calc_block_collision(blk(:, :, :, :, 3), blk(:, :, :, :, 4), blk(:, :, :, :, 1), &
                                      blk(:, :, :, :, 2), nen, en, spindeg, nm, nx, Itot, Ispec)
            call write_spectrum('gw_Scat', iter, Ispec, nen, en, nx, NB, NS, Lx, (/1.0d0, 1.0d0/), .false.)
        end do

        do i = 1, 3
            call qt_check(qt_device_free(Sd(i)), 'qt_device_free')
            call qt_check(qt_device_free(Gd(i)), 'qt_device_free')
            call qt_check(qt_device_free(Pd(i)), 'qt_device_free')
        end do
        call qt_check(qt_device_free(dV(1)), 'qt_device_free')
        call qt_check(qt_device_free(dV(2)), 'qt_device_free')
        call qt_check(qt_ham_destroy(ham), 'qt_ham_destroy')
    end subroutine gw_ephoton_3d_ijs

    ! W^R = V (I - P^R V)^-1, W^<,> = W^R P^<,> W^R^H for nw frequencies in one
    ! call (the reference calls it per energy, gw_mod.f90:592-599); nbc = 0
    ! (closed system) is the only boundary treatment provided
    subroutine green_rgf_calc_w(nbc, nm, nx, Vii, V1i, p_lesser, p_greater, p_retarded, &
                                w_lesser, w_greater, w_retarded, nw)
        integer, intent(in) :: nbc, nm, nx
        integer, intent(in), optional :: nw
        complex(8), intent(in), dimension(nm, nm, nx), target :: Vii, V1i
        complex(8), intent(in), dimension(nm, nm, nx, *), target :: p_lesser, p_greater, p_retarded
        complex(8), intent(inout), dimension(nm, nm, nx, *), target :: w_lesser, w_greater, w_retarded
        integer(c_int) :: rc
        integer :: n
        if (nbc /= 0) then
            write (*, *) 'green_rgf_calc_w: only nbc = 0 (closed system) is provided'
            error stop 1
        end if
        n = 1
        if (present(nw)) n = nw
        call qt_ensure_init()
        rc = qt_gw_wsolve_batched(int(nm, c_int32_t), int(nx, c_int32_t), int(n, c_int32_t), c_loc(Vii), &
                                  c_loc(V1i), c_loc(p_retarded), c_loc(p_lesser), c_loc(p_greater), &
                                  c_loc(w_retarded), c_loc(w_lesser), c_loc(w_greater), 0_c_int32_t)
        if (rc == QT_ERR_SINGULAR) then
            write (*, *) 'green_rgf_calc_w: singular block (zeroed)'
        else
            call qt_check(rc, 'green_rgf_calc_w')
        end if
    end subroutine green_rgf_calc_w

end module gw_mod
