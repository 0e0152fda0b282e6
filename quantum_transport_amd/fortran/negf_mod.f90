! negf_mod.f90 -- ballistic NEGF driver on the MI355X.
!
! negf_solve keeps the reference argument list (negf_mod.f90:41-58 there).
! Instead of one solver call per (E, k) point it hands all local E x k
! points to the device in batched calls of qt_rgf_solve (forward variant,
! the arithmetic of rgf_mod's rgf_variableblock_forward) and keeps only the
! diagonals it writes: ldos / ndos / pdos with coefficients (1,-2), (1,1),
! (1,-1) (negf_mod.f90:141-149 there).  It also writes what the reference's
! ballistic driver computes but drops: the spectral currents at both
! contacts per energy ('trL0000.dat', 'trR0000.dat') and their energy
! integrals in A (appended to 'Id.dat', spin degeneracy 1, single-rank runs),
! with the writers the GW driver calls (gw_mod.f90:409-428 there).
module negf_mod
    use iso_c_binding
    use qt_capi
    use matrix_c, only: type_matrix_complex
    use qt_rgf_host, only: lists_identity
    use Output, only: write_spectrum_from_diagonals, write_transmission_spectrum_k, write_total_current
    implicit none
    private
    integer, parameter :: dp = 8
    integer, parameter :: CHUNK = 256   ! points per device call (bounds host memory)
    public :: negf_solve

contains

    subroutine negf_solve(nx, nen, nk, emin, emax, Hii, H1i, Sii, temp, mu, &
                          comm_size, comm_rank, local_NE, first_local_energy, nbnd, nslab, Lx)
        integer, intent(in) :: nx, nen, nk, nbnd, nslab
        type(type_matrix_complex), intent(in), dimension(nx, nk), target :: Hii, Sii
        type(type_matrix_complex), intent(in), dimension(nx + 1, nk), target :: H1i
        integer(kind=4), intent(in) :: comm_size, comm_rank, local_NE, first_local_energy
        real(dp), intent(in), dimension(2) :: temp, mu
        real(dp), intent(in) :: Lx
        real(dp) :: emin, emax

        real(dp) :: En(nen), dE, local_energies(local_NE)
        integer(c_int32_t), target :: nm(nx)
        type(c_ptr), target :: hp(nx*nk), cp((nx + 1)*nk), sp(nx*nk)
        type(c_ptr) :: ham, sptr
        type(qt_rgf_params) :: p
        type(qt_rgf_outputs) :: o
        complex(dp), allocatable, target :: gr(:, :), gl(:, :), gg(:, :)
        complex(dp), allocatable :: DR(:, :, :), DL(:, :, :), DG(:, :, :)
        real(dp), allocatable :: TL(:, :), TR_(:, :)
        real(c_double), allocatable, target :: tr(:), tre(:), pe(:)
        integer(c_int32_t), allocatable, target :: pk(:), info(:)
        integer :: ie, ik, s, nd, npts, p0, nb, ip, i
        integer(c_int) :: rc
        logical :: identS

        call qt_ensure_init()
        nd = 0
        do s = 1, nx
            nm(s) = size(Hii(s, 1)%m, 1)
            nd = nd + nm(s)
        end do
        identS = .true.
        do ik = 1, nk
            do s = 1, nx
                hp((ik - 1)*nx + s) = c_loc(Hii(s, ik)%m)
                sp((ik - 1)*nx + s) = c_loc(Sii(s, ik)%m)
            end do
            do s = 1, nx + 1
                cp((ik - 1)*(nx + 1) + s) = c_loc(H1i(s, ik)%m)
            end do
            if (.not. lists_identity(Sii(:, ik))) identS = .false.
        end do
        sptr = c_null_ptr
        if (.not. identS) sptr = c_loc(sp)
        call qt_check(qt_ham_create(int(nx, c_int32_t), nm, int(nk, c_int32_t), c_loc(hp), c_loc(cp), &
                                    sptr, ham), 'qt_ham_create')

        dE = (emax - emin)/dble(nen - 1)
        do ie = 1, nen
            En(ie) = emin + dble(ie - 1)*dE
        end do
        do ie = 1, local_NE
            local_energies(ie) = En(ie + first_local_energy - 1)
        end do

        ! points in the reference loop order: energy outer, k inner
        npts = local_NE*nk
        allocate (pe(npts), pk(npts))
        do ie = 1, local_NE
            do ik = 1, nk
                pe((ie - 1)*nk + ik) = local_energies(ie)
                pk((ie - 1)*nk + ik) = ik - 1
            end do
        end do
        allocate (DR(nd, nk, local_NE), DL(nd, nk, local_NE), DG(nd, nk, local_NE))
        allocate (TL(nk, local_NE), TR_(nk, local_NE))
        allocate (gr(nd, CHUNK), gl(nd, CHUNK), gg(nd, CHUNK), tr(CHUNK), tre(CHUNK), info(CHUNK))
        p%variant = QT_RGF_FORWARD
        p%mul = mu(1)
        p%mur = mu(2)
        p%templ = temp(1)
        p%tempr = temp(2)
        o%tr = c_loc(tr)
        o%tre = c_loc(tre)
        o%info = c_loc(info)
        o%gr_diag = c_loc(gr)
        o%gl_diag = c_loc(gl)
        o%gg_diag = c_loc(gg)
        do p0 = 1, npts, CHUNK
            nb = min(CHUNK, npts - p0 + 1)
            rc = qt_rgf_solve(ham, p, int(nb, c_int32_t), pe(p0:), pk(p0:), c_null_ptr, c_null_ptr, o)
            if (rc == QT_ERR_SINGULAR) then
                write (*, *) 'SEVERE warning: singular block inversion (block zeroed)'
            else
                call qt_check(rc, 'qt_rgf_solve')
            end if
            do i = 1, nb
                ip = p0 + i - 1
                ie = (ip - 1)/nk + 1
                ik = ip - (ie - 1)*nk
                DR(:, ik, ie) = gr(:, i)
                DL(:, ik, ie) = gl(:, i)
                DG(:, ik, ie) = gg(:, i)
                TL(ik, ie) = tr(i)
                TR_(ik, ie) = tre(i)
                if (iand(info(i), QT_INFO_LEAD_NMAX) /= 0) &
                    write (*, *) 'SEVERE warning: nmax reached in sancho!!! E =', pe(ip)
            end do
        end do
        call qt_check(qt_ham_destroy(ham), 'qt_ham_destroy')

        do i = 0, comm_size - 1
            if (i == comm_rank) then
                call write_spectrum_from_diagonals('ldos', 0, DR, local_NE, local_energies, nk, nx, nbnd, &
                                                   nslab, Lx, [1.0d0, -2.0d0], comm_rank /= 0)
                call write_spectrum_from_diagonals('ndos', 0, DL, local_NE, local_energies, nk, nx, nbnd, &
                                                   nslab, Lx, [1.0d0, 1.0d0], comm_rank /= 0)
                call write_spectrum_from_diagonals('pdos', 0, DG, local_NE, local_energies, nk, nx, nbnd, &
                                                   nslab, Lx, [1.0d0, -1.0d0], comm_rank /= 0)
                ! 'trL' holds tr and 'trR' tre, the mapping of gw_mod.f90:409-412
                ! there (tr is evaluated at the last contact, tre at the first)
                call write_transmission_spectrum_k('trL', 0, TL, local_NE, local_energies, nk, comm_rank /= 0)
                call write_transmission_spectrum_k('trR', 0, TR_, local_NE, local_energies, nk, comm_rank /= 0)
            end if
        end do
        ! the integral needs every energy: written when one rank holds them all
        if (comm_size == 1) call write_total_current('Id.dat', 0, sum(TL), sum(TR_), dE, 1d0, nk)
    end subroutine negf_solve

end module negf_mod
