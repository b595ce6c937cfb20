! Test driver for the Fortran drop-in boundary of the coh family: the time
! loop of src/main2coh.f95:278-300 (derive2coh, ab2coh, history rotation) with
! the reference's COMMON /comm/ layout (src/main2coh.f95:29), calling the
! reference's routine names (tara_amd/fortran/coh_dropin.f90).
!
! usage: drive_coh IN OUT NSTEPS DT MU CS
!   IN / OUT: int32 Nx, int32 Ny, int32 nfields (3), then nk, ukx, uky as
!   complex(8) (Nh,Ny), column-major.  Lx = Ly = 2 pi.  History starts at zero.
program drive_coh
  implicit none
  real(kind=8), parameter :: pi = 3.14159265358979323846d0
  integer(kind=4) :: Nx, Ny, Nh, nf, t, nsteps, u
  real(kind=8) :: time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8), allocatable, dimension(:,:) :: nk, ukx, uky, nk_new, ukx_new, uky_new
  complex(kind=8), allocatable, dimension(:,:) :: dn_old, dux_old, duy_old, dn_new, dux_new, duy_new
  character(len=512) :: fin, fout, arg
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) dt
  call get_command_argument(5, arg); read(arg, *) mux
  call get_command_argument(6, arg); read(arg, *) CS
  muy = mux
  ms = 1.0d0
  Lx = 2.0d0*pi
  Ly = 2.0d0*pi
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) Nx, Ny, nf
  Nh = Nx/2 + 1
  allocate(nk(Nh,Ny), ukx(Nh,Ny), uky(Nh,Ny), nk_new(Nh,Ny), ukx_new(Nh,Ny), uky_new(Nh,Ny))
  allocate(dn_old(Nh,Ny), dux_old(Nh,Ny), duy_old(Nh,Ny), dn_new(Nh,Ny), dux_new(Nh,Ny), duy_new(Nh,Ny))
  read(u) nk, ukx, uky
  close(u)
  dn_old = (0.0d0, 0.0d0); dux_old = (0.0d0, 0.0d0); duy_old = (0.0d0, 0.0d0)
  time = 0.0d0
  do t = 1, nsteps
    call derive2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, dn_old, dux_old, duy_old, dn_new, dux_new, duy_new)
    call ab2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, nk_new, ukx_new, uky_new, &
                dn_old, dux_old, duy_old, dn_new, dux_new, duy_new)
    dn_old = dn_new; dux_old = dux_new; duy_old = duy_new
    nk = nk_new; ukx = ukx_new; uky = uky_new
    time = time + dt
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Nx, Ny, nf
  write(u) nk, ukx, uky
  close(u)
end program drive_coh
