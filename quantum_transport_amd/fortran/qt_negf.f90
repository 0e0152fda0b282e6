! qt_negf.f90 -- NEGF driver program: reads the namelist /input/ from the
! file 'input' and the Hamiltonian from 'ham_dat' (calc_type = 'w90') or from
! the COO file 'h_dat' (any other calc_type: 0-based real entries, slabs of
! nb orbitals, as main.f90:109-115 there), solves the ballistic transport
! problem on the MI355X and writes ldos/ndos/pdos0000.dat.  Namelist
! variables and defaults follow the reference driver (main.f90:49,75-86
! there); k-points are all zero, one per nk (one for the COO route, whose
! blocks carry no k dependence; each slab is then one output cell, Lx = 1).
program qt_negf
    use omp_lib
    use negf_mod, only: negf_solve
    use matrix_c, only: type_matrix_complex, free
    use deviceHam_mod, only: devH_build_fromWannierFile, devH_build_fromCOOfile
    implicit none
    type(type_matrix_complex), allocatable, dimension(:, :) :: Hii, H1i, Sii
    integer :: nx, ns, nen, nk, nb, rc, fu, nomp
    real(8), dimension(2) :: temp, mu
    real(8) :: emin, emax, Lx, t0, t1
    real(8), allocatable :: k(:, :)
    character(len=10) :: calc_type
    integer(kind=4) :: comm_size, comm_rank, local_NE, first_local_energy
    namelist /input/ nx, ns, temp, mu, nk, nomp, nen, emin, emax, nb, calc_type

    nx = 5; ns = 3; nb = 10; temp = 300d0; mu = 0d0; nk = 1; nen = 100
    emin = -10d0; emax = 5d0; nomp = 4; calc_type = 'w90'
    open (newunit=fu, file='input', action='read', iostat=rc)
    if (rc == 0) then
        read (nml=input, unit=fu, iostat=rc)
        close (fu)
    else
        write (*, *) 'Warn: input file input does not exist'
    end if
    comm_size = 1
    comm_rank = 0
    local_NE = nen/comm_size
    first_local_energy = local_NE*comm_rank + 1
    if (calc_type == 'w90') then
        allocate (k(2, nk))
        k = 0d0
        call devH_build_fromWannierFile('ham_dat', Hii, H1i, Sii, nx, ns, nb, nk, k, Lx)
    else
        call devH_build_fromCOOfile('h_dat', Hii, H1i, Sii, (/0, 0/), (/nb, nb/), nx, &
                                    use0index=.true., iscomplex=.false., threshold=1d-10, blocksize=nb)
        nk = 1
        ns = 1
        Lx = 1d0
    end if
    t0 = omp_get_wtime()
    call negf_solve(nx, nen, nk, emin, emax, Hii, H1i, Sii, temp, mu, comm_size, comm_rank, &
                    local_NE, first_local_energy, nb, ns, Lx)
    t1 = omp_get_wtime()
    write (*, *) 'Total Work took seconds', t1 - t0
    call free(Hii)
    call free(H1i)
    call free(Sii)
end program qt_negf
