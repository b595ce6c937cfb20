! Drop-in replacements for the reference's screened-vorticity routines with
! their names and argument lists (src/derive2css.f95:6, src/ab2css.f95:6,
! src/rk42css.f95:6).  lambda comes from tara_css_lambda (the reference reads an
! uninitialised local; its driver sets 10, src/main2css.f95:87).

subroutine derive2css(Nx, Ny, Nh, alpha, nu, time, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: alpha, nu, time
  complex(kind=8) :: omegak(Nh, Ny), omegak_new(Nh, Ny), dt_omegak_old(Nh, Ny), dt_omegak_new(Nh, Ny)
  call tara_check(tara_derive2css(Nx, Ny, Nh, alpha, nu, tara_css_lambda, omegak, dt_omegak_new), 'derive2css')
end subroutine derive2css

subroutine ab2css(Nx, Ny, Nh, alpha, nu, time, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: alpha, nu, time, dt
  complex(kind=8) :: omegak(Nh, Ny), omegak_new(Nh, Ny), dt_omegak_old(Nh, Ny), dt_omegak_new(Nh, Ny)
  call tara_check(tara_ab2css(Nx, Ny, Nh, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new), 'ab2css')
end subroutine ab2css

! One classic RK4 step (the scheme the reference routine intends).
subroutine rk42css(Nx, Ny, Nh, alpha, nu, time, dt, omegak, omegak_new, dt_omegak_old, dt_omegak_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: alpha, nu, time, dt
  complex(kind=8) :: omegak(Nh, Ny), omegak_new(Nh, Ny), dt_omegak_old(Nh, Ny), dt_omegak_new(Nh, Ny)
  call tara_check(tara_rk42css(Nx, Ny, Nh, alpha, nu, tara_css_lambda, dt, omegak, omegak_new), 'rk42css')
end subroutine rk42css
