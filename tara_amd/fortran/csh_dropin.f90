! Drop-in replacements for the reference's 2D incompressible routines, with
! their names and argument lists (src/derive2csh.f95:3, src/ab2csh.f95:3), so a
! driver such as src/main2csh.f95 links against these instead of the FFTW
! versions.  Arguments the reference does not use (time, omegak_new and
! dt_omegak_old in derive2csh; nu and time in ab2csh) are accepted and ignored.

subroutine derive2csh(Nx, Ny, Nh, nu, time, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: nu, time
  complex(kind=8) :: omegak(Nh, Ny), omegak_new(Nh, Ny), dt_omegak_old(Nh, Ny), dt_omegak_new(Nh, Ny)
  call tara_check(tara_derive2csh(Nx, Ny, Nh, nu, omegak, dt_omegak_new), 'derive2csh')
end subroutine derive2csh

subroutine ab2csh(Nx, Ny, Nh, nu, time, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: nu, time, dt
  complex(kind=8) :: omegak(Nh, Ny), omegak_new(Nh, Ny), dt_omegak_old(Nh, Ny), dt_omegak_new(Nh, Ny)
  call tara_check(tara_ab2csh(Nx, Ny, Nh, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new), 'ab2csh')
end subroutine ab2csh
