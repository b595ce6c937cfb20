! Test driver for the Fortran drop-in boundary of the 3D MHD path: the loop of
! src/main3com.f95:536-617 (derive3com, ab3com, rotation) with the reference's
! COMMON /comm/ layout (src/main3com.f95:56).
!
! usage: drive_mhd IN OUT NSTEPS DT CS [MU ETA]
!   IN / OUT: int32 Nx, Ny, Nz, nfields (8), then the 8 state fields as
!   complex(8) (Nh,Ny,Nz).  Lx = Ly = Lz = 2 pi.  With MU and ETA given, the
!   diffusion takes them from COMMON; otherwise 1/450 as the reference.
program drive_mhd
  use tara_hip_mod, only: tara_mhd_diffusion_from_common
  implicit none
  real(kind=8), parameter :: pi = 3.14159265358979323846d0
  integer(kind=4) :: Nx, Ny, Nz, Nh, nf, t, nsteps, u, f, iret, thread_num
  real(kind=8) :: time, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  complex(kind=8), allocatable :: s(:,:,:,:), sn(:,:,:,:), fo(:,:,:,:), fn(:,:,:,:)
  character(len=512) :: fin, fout, arg
  common/comm/iret, thread_num, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) dt
  call get_command_argument(5, arg); read(arg, *) CS
  mu = 0.0d0; eta = 0.0d0
  if (command_argument_count() >= 7) then
    call get_command_argument(6, arg); read(arg, *) mu
    call get_command_argument(7, arg); read(arg, *) eta
    tara_mhd_diffusion_from_common = .true.
  end if
  Lx = 2.0d0*pi; Ly = 2.0d0*pi; Lz = 2.0d0*pi
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) Nx, Ny, Nz, nf
  Nh = Nx/2 + 1
  dx = Lx/dfloat(Nx); dy = Ly/dfloat(Ny); dz = Lz/dfloat(Nz)
  allocate(s(Nh,Ny,Nz,8), sn(Nh,Ny,Nz,8), fo(Nh,Ny,Nz,8), fn(Nh,Ny,Nz,8))
  read(u) s
  close(u)
  fo = (0.0d0, 0.0d0)
  time = 0.0d0
  do t = 1, nsteps
    call derive3com(Nx, Ny, Nz, Nh, pi, time, s(:,:,:,1), s(:,:,:,2), s(:,:,:,3), s(:,:,:,4), s(:,:,:,5), &
                    s(:,:,:,6), s(:,:,:,7), s(:,:,:,8), &
                    fo(:,:,:,1), fn(:,:,:,1), fo(:,:,:,2), fn(:,:,:,2), fo(:,:,:,3), fn(:,:,:,3), &
                    fo(:,:,:,4), fn(:,:,:,4), fo(:,:,:,5), fn(:,:,:,5), fo(:,:,:,6), fn(:,:,:,6), &
                    fo(:,:,:,7), fn(:,:,:,7), fo(:,:,:,8), fn(:,:,:,8))
    call ab3com(Nx, Ny, Nz, Nh, pi, time, s(:,:,:,1), s(:,:,:,2), s(:,:,:,3), s(:,:,:,4), s(:,:,:,5), &
                s(:,:,:,6), s(:,:,:,7), s(:,:,:,8), &
                sn(:,:,:,1), sn(:,:,:,2), sn(:,:,:,3), sn(:,:,:,4), sn(:,:,:,5), sn(:,:,:,6), sn(:,:,:,7), &
                sn(:,:,:,8), &
                fo(:,:,:,1), fn(:,:,:,1), fo(:,:,:,2), fn(:,:,:,2), fo(:,:,:,3), fn(:,:,:,3), &
                fo(:,:,:,4), fn(:,:,:,4), fo(:,:,:,5), fn(:,:,:,5), fo(:,:,:,6), fn(:,:,:,6), &
                fo(:,:,:,7), fn(:,:,:,7), fo(:,:,:,8), fn(:,:,:,8))
    do f = 1, 8
      fo(:,:,:,f) = fn(:,:,:,f)
      s(:,:,:,f) = sn(:,:,:,f)
    end do
    time = time + dt
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Nx, Ny, Nz, nf
  write(u) s
  close(u)
end program drive_mhd
