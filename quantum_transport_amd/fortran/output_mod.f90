! output_mod.f90 -- spectrum writers of the NEGF driver (module Output).
!
! Record layout of the reference writer (output_mod.f90:14-61 there): one
! line per cell and energy, '(4E18.4)' = x, E, coeff(1)*Re tr, coeff(2)*Im tr
! with tr the k-averaged trace of the cell's diagonal block, and an empty
! record after each energy.  File name: dataset // '%4.4i' // '.dat'.
module Output
    use matrix_c, only: type_matrix_complex
    implicit none
    private
    integer, parameter :: dp = 8
    public :: write_spectrum_summed_over_k, write_spectrum_from_diagonals, write_transmission_spectrum_k, &
              write_total_current, write_spectrum, write_current_spectrum_summed_over_kz, &
              write_current_summed_over_k
    ! constants of the reference's current formula (static.f90:10,14, fft_mod.f90:19 there)
    real(dp), parameter :: hbar = 1.0546d-34, e0 = 1.6022d-19, tpi = 3.14159265359d0*2.0d0

contains

    subroutine open_spectrum(dataset, i, append)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i
        logical, intent(in) :: append
        character(len=4) :: tag
        write (tag, '(I4.4)') i
        if (append) then
            open (unit=11, file=trim(dataset)//tag//'.dat', status='unknown', position='append')
        else
            open (unit=11, file=trim(dataset)//tag//'.dat', status='unknown')
        end if
    end subroutine open_spectrum

    ! G(length, nen, nk): diagonal blocks per slab, energy and k-point
    subroutine write_spectrum_summed_over_k(dataset, i, G, nen, en, nk, length, NB, NS, Lx, coeff, append)
        character(len=*), intent(in) :: dataset
        type(type_matrix_complex), intent(in), dimension(length, nen, nk) :: G
        integer, intent(in) :: i, nen, length, NB, NS, nk
        real(dp), intent(in) :: Lx, en(nen), coeff(2)
        logical, intent(in), optional :: append
        logical :: app
        integer :: ie, j, c, ik, b, cell
        complex(dp) :: t
        app = .false.
        if (present(append)) app = append
        call open_spectrum(dataset, i, app)
        do ie = 1, nen
            cell = 0
            do j = 1, length
                do c = 1, NS
                    cell = cell + 1
                    t = (0d0, 0d0)
                    do ik = 1, nk
                        do b = 1, NB
                            t = t + G(j, ie, ik)%m((c - 1)*NB + b, (c - 1)*NB + b)
                        end do
                    end do
                    t = t/dble(nk)
                    write (11, '(4E18.4)') dble(cell)*Lx, en(ie), dble(t)*coeff(1), aimag(t)*coeff(2)
                end do
            end do
            write (11, '(A)') ''
        end do
        close (11)
    end subroutine write_spectrum_summed_over_k

    ! D(nd, nk, nen): per energy and k-point the concatenated diagonals of
    ! the length slabs (nd = length*NS*NB)
    subroutine write_spectrum_from_diagonals(dataset, i, D, nen, en, nk, length, NB, NS, Lx, coeff, append)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i, nen, length, NB, NS, nk
        complex(dp), intent(in) :: D(length*NS*NB, nk, nen)
        real(dp), intent(in) :: Lx, en(nen), coeff(2)
        logical, intent(in), optional :: append
        logical :: app
        integer :: ie, cell, ik, o
        complex(dp) :: t
        app = .false.
        if (present(append)) app = append
        call open_spectrum(dataset, i, app)
        do ie = 1, nen
            do cell = 1, length*NS
                o = (cell - 1)*NB
                t = (0d0, 0d0)
                do ik = 1, nk
                    t = t + sum(D(o + 1:o + NB, ik, ie))
                end do
                t = t/dble(nk)
                write (11, '(4E18.4)') dble(cell)*Lx, en(ie), dble(t)*coeff(1), aimag(t)*coeff(2)
            end do
            write (11, '(A)') ''
        end do
        close (11)
    end subroutine write_spectrum_from_diagonals

    ! Spectral current per energy, the file the GW driver names 'gw_trL' /
    ! 'gw_trR' and never defines (gw_mod.f90:409-412 there): one record
    ! '(2E18.4)' per energy, E and the k-average of T(E, k).
    subroutine write_transmission_spectrum_k(dataset, i, T, nen, en, nk, append)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i, nen, nk
        real(dp), intent(in) :: T(nk, nen), en(nen)
        logical, intent(in), optional :: append
        logical :: app
        integer :: ie
        app = .false.
        if (present(append)) app = append
        call open_spectrum(dataset, i, app)
        do ie = 1, nen
            write (11, '(2E18.4)') en(ie), sum(T(:, ie))/dble(nk)
        end do
        close (11)
    end subroutine write_transmission_spectrum_k

    ! Energy-integrated terminal currents in A, the record gw_mod.f90:424-428
    ! there appends to 'Id_iteration.dat': '(I4,2E16.6)' iteration,
    ! sum(trL) dE e0 / (2 pi hbar) e0 spindeg / nk, the same for trR.
    subroutine write_total_current(fname, iter, sum_tr, sum_tre, dE, spindeg, nk)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: iter, nk
        real(dp), intent(in) :: sum_tr, sum_tre, dE, spindeg
        open (unit=101, file=trim(fname), status='unknown', position='append')
        write (101, '(I4,2E16.6)') iter, sum_tr*dE*e0/tpi/hbar*e0*spindeg/dble(nk), &
            sum_tre*dE*e0/tpi/hbar*e0*spindeg/dble(nk)
        close (101)
    end subroutine write_total_current

    ! The spectrum of plain (nb ns, nb ns, length, nen) blocks, the writer the
    ! GW driver calls for its collision integrals ('gw_Scat', 'eph_Scat',
    ! gw_mod.f90:803, :934 there) and never defines: the records of
    ! write_spectrum_summed_over_k for one k point.
    subroutine write_spectrum(dataset, i, G, nen, en, length, NB, NS, Lx, coeff, append)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i, nen, length, NB, NS
        complex(dp), intent(in) :: G(NB*NS, NB*NS, length, nen)
        real(dp), intent(in) :: Lx, en(nen), coeff(2)
        logical, intent(in), optional :: append
        logical :: app
        integer :: ie, j, c, b, cell
        complex(dp) :: t
        app = .false.
        if (present(append)) app = append
        call open_spectrum(dataset, i, app)
        do ie = 1, nen
            cell = 0
            do j = 1, length
                do c = 1, NS
                    cell = cell + 1
                    t = (0d0, 0d0)
                    do b = 1, NB
                        t = t + G((c - 1)*NB + b, (c - 1)*NB + b, j, ie)
                    end do
                    write (11, '(4E18.4)') dble(cell)*Lx, en(ie), dble(t)*coeff(1), aimag(t)*coeff(2)
                end do
            end do
            write (11, '(A)') ''
        end do
        close (11)
    end subroutine write_spectrum

    ! Bond-current spectrum, the file the GW driver names 'gw_Jdens'
    ! (gw_mod.f90:405-406 there) and never defines: one record '(3E18.4)'
    ! per cell bond and energy, x of the bond's cell, E and the k-average of
    ! the trace of the bond's current matrix J(:, :, cell, ie, ik); an empty
    ! record after each energy.
    subroutine write_current_spectrum_summed_over_kz(dataset, i, J, nen, en, length, NB, Lx, nk, append)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i, nen, length, NB, nk
        real(dp), intent(in) :: J(NB, NB, length, nen, nk), en(nen), Lx
        logical, intent(in), optional :: append
        logical :: app
        integer :: ie, cell, ik, b
        real(dp) :: t
        app = .false.
        if (present(append)) app = append
        call open_spectrum(dataset, i, app)
        do ie = 1, nen
            do cell = 1, length
                t = 0d0
                do ik = 1, nk
                    do b = 1, NB
                        t = t + J(b, b, cell, ie, ik)
                    end do
                end do
                write (11, '(3E18.4)') dble(cell)*Lx, en(ie), t/dble(nk)
            end do
            write (11, '(A)') ''
        end do
        close (11)
    end subroutine write_current_spectrum_summed_over_kz

    ! Energy-integrated bond currents, the files the GW driver names 'gw_I' and
    ! 'gw_EI' (gw_mod.f90:429-430 there) and never defines: one record
    ! '(2E18.4)' per cell bond, x of the bond's cell and the k-average of the
    ! trace of Cur(:, :, cell, ik).
    subroutine write_current_summed_over_k(dataset, i, Cur, length, NB, Lx, nk)
        character(len=*), intent(in) :: dataset
        integer, intent(in) :: i, length, NB, nk
        real(dp), intent(in) :: Cur(NB, NB, length, nk), Lx
        integer :: cell, ik, b
        real(dp) :: t
        call open_spectrum(dataset, i, .false.)
        do cell = 1, length
            t = 0d0
            do ik = 1, nk
                do b = 1, NB
                    t = t + Cur(b, b, cell, ik)
                end do
            end do
            write (11, '(2E18.4)') dble(cell)*Lx, t/dble(nk)
        end do
        close (11)
    end subroutine write_current_summed_over_k

end module Output
