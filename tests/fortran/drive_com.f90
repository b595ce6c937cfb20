! Test driver for the Fortran drop-in boundary of the com family: the time
! loop of src/main2com.f95:364-380 (derive2com, ab2com, history rotation) with
! the reference's COMMON /comm/ layout (src/main2com.f95:43).
!
! usage: drive_com IN OUT NSTEPS DT MU ETA CS
!   IN / OUT: int32 Nx, int32 Ny, int32 nfields (6), then rho, rho ux, rho uy,
!   E, Bx, By as complex(8) (Nh,Ny).  Lx = Ly = 2 pi, forcing off (A = 0).
program drive_com
  implicit none
  real(kind=8), parameter :: pi = 3.14159265358979323846d0
  integer(kind=4) :: Nx, Ny, Nh, nf, t, nsteps, u, f
  real(kind=8) :: time, dt, Lx, Ly, spheat, mu, ms, CS, mu_0, eta, kf, A
  complex(kind=8), allocatable :: s(:,:,:), sn(:,:,:), fo(:,:,:), fn(:,:,:)
  character(len=512) :: fin, fout, arg
  common/comm/Lx, Ly, spheat, mu, ms, CS, mu_0, eta, dt, kf, A
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) dt
  call get_command_argument(5, arg); read(arg, *) mu
  call get_command_argument(6, arg); read(arg, *) eta
  call get_command_argument(7, arg); read(arg, *) CS
  Lx = 2.0d0*pi; Ly = 2.0d0*pi
  spheat = 5.0d0/3.0d0; ms = 1.0d0; mu_0 = 1.0d0; kf = 0.0d0; A = 0.0d0
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) Nx, Ny, nf
  Nh = Nx/2 + 1
  allocate(s(Nh,Ny,6), sn(Nh,Ny,6), fo(Nh,Ny,6), fn(Nh,Ny,6))
  read(u) s
  close(u)
  fo = (0.0d0, 0.0d0)
  time = 0.0d0
  do t = 1, nsteps
    call derive2com(Nx, Ny, Nh, pi, time, s(:,:,1), s(:,:,2), s(:,:,3), s(:,:,4), s(:,:,5), s(:,:,6), &
                    fo(:,:,1), fn(:,:,1), fo(:,:,2), fn(:,:,2), fo(:,:,3), fn(:,:,3), &
                    fo(:,:,4), fn(:,:,4), fo(:,:,5), fn(:,:,5), fo(:,:,6), fn(:,:,6))
    call ab2com(Nx, Ny, Nh, pi, time, s(:,:,1), s(:,:,2), s(:,:,3), s(:,:,4), s(:,:,5), s(:,:,6), &
                sn(:,:,1), sn(:,:,2), sn(:,:,3), sn(:,:,4), sn(:,:,5), sn(:,:,6), &
                fo(:,:,1), fn(:,:,1), fo(:,:,2), fn(:,:,2), fo(:,:,3), fn(:,:,3), &
                fo(:,:,4), fn(:,:,4), fo(:,:,5), fn(:,:,5), fo(:,:,6), fn(:,:,6))
    do f = 1, 6
      fo(:,:,f) = fn(:,:,f)
      s(:,:,f) = sn(:,:,f)
    end do
    time = time + dt
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Nx, Ny, nf
  write(u) s
  close(u)
end program drive_com
