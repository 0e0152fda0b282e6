! point_check.f90 -- test driver of the per-point Fortran entry points.
!
! Reads one device and one energy from 'point_in.bin' (stream, native
! byte order):
!   int32 nx, nm(nx), mode;  float64 En, mul, mur, TEMPl, TEMPr;
!   complex128 Hii(s) (nm(s) x nm(s), s = 1..nx), H1i(s) (s = 1..nx+1:
!   nm(1) x nm(1), then nm(s) x nm(s-1), then nm(nx) x nm(nx)), all
!   column-major.
! mode 0 calls rgf_mod::rgf_variableblock_forward and _backward,
! cuda_rgf_mod::cuda_rgf_constblocksize and selfenergy_mod::sancho on the
! left lead (H00 = Hii(1), H10 = H1i(1)); mode 1 makes mul a non-uniform
! matrix, which the device path refuses.  Every output block is filled
! with the sentinel (1234.5, -678.25) before each call.  Writes
! 'point_out.bin': for
! each of forward, backward: G_r, G_lesser, G_greater, Jdens, Gl, Gln
! blocks (s = 1..nx), tr, tre; for constblocksize: G_r, G_lesser,
! G_greater, Jdens, tr, tre; then sancho's G00, GBB.
program point_check
    use matrix_c, only: type_matrix_complex, malloc
    use rgf_mod, only: rgf_variableblock_forward, rgf_variableblock_backward
    use cuda_rgf_mod, only: cuda_rgf_init, cuda_rgf_constblocksize, cuda_rgf_finish
    use selfenergy_mod, only: sancho
    implicit none
    integer, parameter :: dp = 8
    integer :: nx, mode, s, i, u
    integer, allocatable :: nm(:)
    real(dp) :: En, vmul, vmur, vtl, vtr, tr, tre
    real(dp), allocatable :: mul(:, :), mur(:, :), TEMPl(:, :), TEMPr(:, :)
    type(type_matrix_complex), allocatable :: Hii(:), H1i(:), Sii(:), sl(:), sr(:)
    type(type_matrix_complex), allocatable :: G_r(:), G_lesser(:), G_greater(:), Jdens(:), Gl(:), Gln(:)
    complex(dp), allocatable :: S00(:, :), G00(:, :), GBB(:, :)

    open (newunit=u, file='point_in.bin', access='stream', form='unformatted', status='old')
    read (u) nx
    allocate (nm(nx))
    read (u) nm, mode, En, vmul, vmur, vtl, vtr
    allocate (Hii(nx), H1i(nx + 1), Sii(nx), sl(nx), sr(nx))
    allocate (G_r(nx), G_lesser(nx), G_greater(nx), Jdens(nx), Gl(nx), Gln(nx))
    do s = 1, nx
        call malloc(Hii(s), nm(s))
        read (u) Hii(s)%m
    end do
    call malloc(H1i(1), nm(1))
    read (u) H1i(1)%m
    do s = 2, nx
        call malloc(H1i(s), nm(s), nm(s - 1))
        read (u) H1i(s)%m
    end do
    call malloc(H1i(nx + 1), nm(nx))
    read (u) H1i(nx + 1)%m
    close (u)
    do s = 1, nx
        call malloc(Sii(s), nm(s))
        do i = 1, nm(s)
            Sii(s)%m(i, i) = (1d0, 0d0)
        end do
        call malloc(sl(s), nm(s))
        call malloc(sr(s), nm(s))
    end do
    allocate (mul(nm(1), nm(1)), mur(nm(nx), nm(nx)), TEMPl(nm(1), nm(1)), TEMPr(nm(nx), nm(nx)))
    mul = vmul
    mur = vmur
    TEMPl = vtl
    TEMPr = vtr
    if (mode == 1) mul(1, 1) = vmul + 0.05d0

    open (newunit=u, file='point_out.bin', access='stream', form='unformatted', status='replace')
    call fill_lists()
    call rgf_variableblock_forward(nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sl, sr, &
                                   G_r, G_lesser, G_greater, Jdens, Gl, Gln, tr, tre)
    call put_lists(6)
    call fill_lists()
    call rgf_variableblock_backward(nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sl, sr, &
                                    G_r, G_lesser, G_greater, Jdens, Gl, Gln, tr, tre)
    call put_lists(6)
    call fill_lists()
    call cuda_rgf_init(nm(1))
    call cuda_rgf_constblocksize(nm(1), nx, En, mul, mur, TEMPl, TEMPr, Hii, H1i, Sii, sl, sr, &
                                 G_r, G_lesser, G_greater, Jdens, tr, tre)
    call cuda_rgf_finish()
    call put_lists(4)
    allocate (S00(nm(1), nm(1)), G00(nm(1), nm(1)), GBB(nm(1), nm(1)))
    S00 = (0d0, 0d0)
    do i = 1, nm(1)
        S00(i, i) = (1d0, 0d0)
    end do
    call sancho(nm(1), En, S00, Hii(1)%m, H1i(1)%m, G00, GBB)
    write (u) G00, GBB
    close (u)

contains

    ! every output block set to SENTINEL before a call: blocks a routine
    ! leaves untouched come back as SENTINEL
    subroutine fill_lists()
        complex(dp), parameter :: SENTINEL = (1234.5d0, -678.25d0)
        do s = 1, nx
            call malloc(G_r(s), nm(s))
            call malloc(G_lesser(s), nm(s))
            call malloc(G_greater(s), nm(s))
            call malloc(Jdens(s), nm(s))
            call malloc(Gl(s), nm(s))
            call malloc(Gln(s), nm(s))
            G_r(s)%m = SENTINEL
            G_lesser(s)%m = SENTINEL
            G_greater(s)%m = SENTINEL
            Jdens(s)%m = SENTINEL
            Gl(s)%m = SENTINEL
            Gln(s)%m = SENTINEL
        end do
    end subroutine fill_lists

    subroutine put_lists(count)
        integer, intent(in) :: count
        integer :: t
        do t = 1, count
            do s = 1, nx
                select case (t)
                case (1)
                    write (u) G_r(s)%m
                case (2)
                    write (u) G_lesser(s)%m
                case (3)
                    write (u) G_greater(s)%m
                case (4)
                    write (u) Jdens(s)%m
                case (5)
                    write (u) Gl(s)%m
                case (6)
                    write (u) Gln(s)%m
                end select
            end do
        end do
        write (u) tr, tre
    end subroutine put_lists

end program point_check
