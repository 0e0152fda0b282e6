! scba_check.f90 -- test driver of gw_mod::gw_ephoton_3d_ijs.
!
! Reads 'scba_in.bin' (stream, native byte order):
!   int32 NB, NS, nx, nen, niter, ndiag;  float64 alpha_mix, spindeg, Lx,
!   temp(2), mu(2), en(nen);  complex128 Hii, H1i, Vii, V1i (nm, nm, nx),
!   nm = NB NS, column-major.
! Calls gw_ephoton_3d_ijs for one process, one k point and no e-photon
! coupling; its files (gw_ldos, gw_Jdens, gw_trL, gw_trR, Id_iteration.dat,
! gw_I, gw_EI, gw_Scat) are the output.
program scba_check
    use gw_mod, only: gw_ephoton_3d_ijs
    implicit none
    integer, parameter :: dp = 8
    integer :: NB, NS, nx, nen, niter, ndiag, nm, u
    real(dp) :: alpha_mix, spindeg, Lx, temp(2), mu(2), pol(3)
    real(dp), allocatable :: en(:)
    complex(dp), allocatable :: Hii(:, :, :, :), H1i(:, :, :, :), Vii(:, :, :, :), V1i(:, :, :, :)
    complex(dp), allocatable :: Pii(:, :, :, :, :), P1i(:, :, :, :, :)

    open (newunit=u, file='scba_in.bin', access='stream', form='unformatted', status='old')
    read (u) NB, NS, nx, nen, niter, ndiag
    nm = NB*NS
    allocate (en(nen))
    read (u) alpha_mix, spindeg, Lx, temp, mu, en
    allocate (Hii(nm, nm, nx, 1), H1i(nm, nm, nx, 1), Vii(nm, nm, nx, 1), V1i(nm, nm, nx, 1))
    read (u) Hii, H1i, Vii, V1i
    close (u)
    allocate (Pii(nm, nm, 3, nx, 1), P1i(nm, nm, 3, nx, 1))
    Pii = (0d0, 0d0)
    P1i = (0d0, 0d0)
    pol = 0d0
    call gw_ephoton_3d_ijs(alpha_mix, niter, NB, NS, nm, nx, 1, 1, ndiag, Lx, nen, en, temp, mu, &
                           Hii, H1i, Vii, V1i, spindeg, Pii, P1i, pol, 0d0, 0d0, .false., 1, 0, nen, 1)
end program scba_check
