! Drop-in replacements for the reference's 2D compressible isothermal hydro
! routines (src/derive2coh.f95:5, src/ab2coh.f95:5, src/euler2coh.f95:5,
! src/rk4nk2coh.f95:6 and twins, src/pcnk2coh.f95:6 and twins), with their
! names, argument lists and COMMON /comm/ layout (src/main2coh.f95:29).

subroutine derive2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, &
                      dn_dt_old, dux_dt_old, duy_dt_old, dn_dt_new, dux_dt_new, duy_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8), target :: nk(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny)
  complex(kind=8) :: dn_dt_old(Nh, Ny), dux_dt_old(Nh, Ny), duy_dt_old(Nh, Ny)
  complex(kind=8), target :: dn_dt_new(Nh, Ny), dux_dt_new(Nh, Ny), duy_dt_new(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  type(tara_coh_params) :: p
  p = tara_coh_params(pi, Lx, Ly, mux, muy, CS, dt)
  call tara_check(tara_derive2coh(Nx, Ny, Nh, p, [c_loc(nk), c_loc(ukx), c_loc(uky)], &
                                  [c_loc(dn_dt_new), c_loc(dux_dt_new), c_loc(duy_dt_new)]), 'derive2coh')
end subroutine derive2coh

subroutine ab2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, nk_new, ukx_new, uky_new, &
                  dn_dt_old, dux_dt_old, duy_dt_old, dn_dt_new, dux_dt_new, duy_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8), target :: nk(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), nk_new(Nh, Ny), ukx_new(Nh, Ny), uky_new(Nh, Ny)
  complex(kind=8), target :: dn_dt_old(Nh, Ny), dux_dt_old(Nh, Ny), duy_dt_old(Nh, Ny)
  complex(kind=8), target :: dn_dt_new(Nh, Ny), dux_dt_new(Nh, Ny), duy_dt_new(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_ab2coh(Nx, Ny, Nh, dt, [c_loc(nk), c_loc(ukx), c_loc(uky)], &
                              [c_loc(nk_new), c_loc(ukx_new), c_loc(uky_new)], &
                              [c_loc(dn_dt_old), c_loc(dux_dt_old), c_loc(duy_dt_old)], &
                              [c_loc(dn_dt_new), c_loc(dux_dt_new), c_loc(duy_dt_new)]), 'ab2coh')
end subroutine ab2coh

subroutine euler2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, nk_new, ukx_new, uky_new, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8), target :: nk(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), nk_new(Nh, Ny), ukx_new(Nh, Ny), uky_new(Nh, Ny)
  complex(kind=8), target :: dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_euler2coh(Nx, Ny, Nh, dt, [c_loc(nk), c_loc(ukx), c_loc(uky)], &
                                 [c_loc(nk_new), c_loc(ukx_new), c_loc(uky_new)], &
                                 [c_loc(dn_dt), c_loc(dux_dt), c_loc(duy_dt)]), 'euler2coh')
end subroutine euler2coh

subroutine rk4_nk2coh(Nx, Ny, Nh, pi, time, nk, nk_new, ukx, uky, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), nk_new(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_rk4coh_field(Nx, Ny, Nh, dt, nk, nk_new, dn_dt), 'rk4_nk2coh')
end subroutine rk4_nk2coh

subroutine rk4_ukx2coh(Nx, Ny, Nh, pi, time, nk, ukx, ukx_new, uky, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), ukx(Nh, Ny), ukx_new(Nh, Ny), uky(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_rk4coh_field(Nx, Ny, Nh, dt, ukx, ukx_new, dux_dt), 'rk4_ukx2coh')
end subroutine rk4_ukx2coh

subroutine rk4_uky2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, uky_new, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), uky_new(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_rk4coh_field(Nx, Ny, Nh, dt, uky, uky_new, duy_dt), 'rk4_uky2coh')
end subroutine rk4_uky2coh

subroutine pc_nk2coh(Nx, Ny, Nh, pi, time, nk, nk_new, ukx, uky, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), nk_new(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_pccoh_field(Nx, Ny, Nh, dt, nk, nk_new, dn_dt), 'pc_nk2coh')
end subroutine pc_nk2coh

subroutine pc_ukx2coh(Nx, Ny, Nh, pi, time, nk, ukx, ukx_new, uky, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), ukx(Nh, Ny), ukx_new(Nh, Ny), uky(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_pccoh_field(Nx, Ny, Nh, dt, ukx, ukx_new, dux_dt), 'pc_ukx2coh')
end subroutine pc_ukx2coh

subroutine pc_uky2coh(Nx, Ny, Nh, pi, time, nk, ukx, uky, uky_new, dn_dt, dux_dt, duy_dt)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mux, muy, ms, CS
  complex(kind=8) :: nk(Nh, Ny), ukx(Nh, Ny), uky(Nh, Ny), uky_new(Nh, Ny), dn_dt(Nh, Ny), dux_dt(Nh, Ny), duy_dt(Nh, Ny)
  common/comm/Lx, Ly, mux, muy, ms, CS, dt
  call tara_check(tara_pccoh_field(Nx, Ny, Nh, dt, uky, uky_new, duy_dt), 'pc_uky2coh')
end subroutine pc_uky2coh
