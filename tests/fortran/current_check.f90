! current_check.f90 -- test driver of gw_mod::calc_block_current,
! gw_mod::calc_block_collision and the Output writers of their results.
!
! Reads 'current_in.bin' (stream, native byte order):
!   int32 nb, ns, nx, nen, nk;  float64 spindeg, Lx, en(nen);
!   complex128 Hii (nm, nm, nx, nk), GL (nm, nm, nx, nen, nk);
!   float64 cur (nm, nm, nx, nen, nk);
!   complex128 Sl, Sg, Gl, Gg (nm, nm, nx, nen), nm = nb ns, column-major.
! Calls calc_block_current per k point (as gw_mod.f90:384 does there) and
! calc_block_collision once, writes 'current_out.bin' with jdens (nb, nb,
! nx ns, nen, nk), tot_cur, tot_ecur (nb, nb, nx ns, nk), Itot (nm, nm, nx),
! Ispec (nm, nm, nx, nen), and the files gw_Jdens0001.dat, gw_I0001.dat,
! gw_EI0001.dat and gw_Scat0001.dat with the driver's calls
! (gw_mod.f90:405-406, :429-430, :803 there).
program current_check
    use gw_mod, only: calc_block_current, calc_block_collision
    use Output, only: write_current_spectrum_summed_over_kz, write_current_summed_over_k, write_spectrum
    implicit none
    integer, parameter :: dp = 8
    integer :: nb, ns, nx, nen, nk, nm, ik, u
    real(dp) :: spindeg, Lx
    real(dp), allocatable :: en(:), cur(:, :, :, :, :), jdens(:, :, :, :, :), tot_cur(:, :, :, :), &
                             tot_ecur(:, :, :, :)
    complex(dp), allocatable :: Hii(:, :, :, :), GL(:, :, :, :, :), Sl(:, :, :, :), Sg(:, :, :, :), &
                                Gl2(:, :, :, :), Gg(:, :, :, :), Itot(:, :, :), Ispec(:, :, :, :)

    open (newunit=u, file='current_in.bin', access='stream', form='unformatted', status='old')
    read (u) nb, ns, nx, nen, nk
    nm = nb*ns
    allocate (en(nen))
    read (u) spindeg, Lx, en
    allocate (Hii(nm, nm, nx, nk), GL(nm, nm, nx, nen, nk), cur(nm, nm, nx, nen, nk))
    allocate (Sl(nm, nm, nx, nen), Sg(nm, nm, nx, nen), Gl2(nm, nm, nx, nen), Gg(nm, nm, nx, nen))
    read (u) Hii, GL, cur, Sl, Sg, Gl2, Gg
    close (u)
    allocate (jdens(nb, nb, nx*ns, nen, nk), tot_cur(nb, nb, nx*ns, nk), tot_ecur(nb, nb, nx*ns, nk))
    allocate (Itot(nm, nm, nx), Ispec(nm, nm, nx, nen))
    do ik = 1, nk
        call calc_block_current(Hii(:, :, :, ik), GL(:, :, :, :, ik), cur(:, :, :, :, ik), nen, en, spindeg, &
                                nb, ns, nm, nx, tot_cur(:, :, :, ik), tot_ecur(:, :, :, ik), jdens(:, :, :, :, ik))
    end do
    call calc_block_collision(Sl, Sg, Gl2, Gg, nen, en, spindeg, nm, nx, Itot, Ispec)
    open (newunit=u, file='current_out.bin', access='stream', form='unformatted', status='replace')
    write (u) jdens, tot_cur, tot_ecur, Itot, Ispec
    close (u)
    call write_current_spectrum_summed_over_kz('gw_Jdens', 1, jdens, nen, en, nx*ns, nb, Lx, nk, .false.)
    call write_current_summed_over_k('gw_I', 1, tot_cur, nx*ns, nb, Lx, nk)
    call write_current_summed_over_k('gw_EI', 1, tot_ecur, nx*ns, nb, Lx, nk)
    call write_spectrum('gw_Scat', 1, Ispec, nen, en, nx, nb, ns, Lx, (/1.0d0/dble(nk), 1.0d0/dble(nk)/), .false.)
end program current_check
