! Drop-in replacements for the reference's 2D compressible MHD routines
! (src/derive2com.f95:5, src/ab2com.f95:6) with their names, argument lists and
! COMMON /comm/ layout (src/main2com.f95:43).

subroutine derive2com(Nx, Ny, Nh, pi, time, rho_k, rho_ux_k, rho_uy_k, E_k, Bx_k, By_k, &
                      d_rho_k_dt_old, d_rho_k_dt_new, d_rho_ux_k_dt_old, d_rho_ux_k_dt_new, &
                      d_rho_uy_k_dt_old, d_rho_uy_k_dt_new, d_E_k_dt_old, d_E_k_dt_new, &
                      d_Bx_k_dt_old, d_Bx_k_dt_new, d_By_k_dt_old, d_By_k_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mu, ms, CS, mu_0, eta, spheat, kf, A
  complex(kind=8), target :: rho_k(Nh, Ny), rho_ux_k(Nh, Ny), rho_uy_k(Nh, Ny), E_k(Nh, Ny), Bx_k(Nh, Ny), By_k(Nh, Ny)
  complex(kind=8) :: d_rho_k_dt_old(Nh, Ny), d_rho_ux_k_dt_old(Nh, Ny), d_rho_uy_k_dt_old(Nh, Ny)
  complex(kind=8) :: d_E_k_dt_old(Nh, Ny), d_Bx_k_dt_old(Nh, Ny), d_By_k_dt_old(Nh, Ny)
  complex(kind=8), target :: d_rho_k_dt_new(Nh, Ny), d_rho_ux_k_dt_new(Nh, Ny), d_rho_uy_k_dt_new(Nh, Ny)
  complex(kind=8), target :: d_E_k_dt_new(Nh, Ny), d_Bx_k_dt_new(Nh, Ny), d_By_k_dt_new(Nh, Ny)
  common/comm/Lx, Ly, spheat, mu, ms, CS, mu_0, eta, dt, kf, A
  type(tara_com_params) :: p
  p = tara_com_params(pi, Lx, Ly, mu, eta, CS, kf, A, dt)
  call tara_check(tara_derive2com(Nx, Ny, Nh, p, &
      [c_loc(rho_k), c_loc(rho_ux_k), c_loc(rho_uy_k), c_loc(E_k), c_loc(Bx_k), c_loc(By_k)], &
      [c_loc(d_rho_k_dt_new), c_loc(d_rho_ux_k_dt_new), c_loc(d_rho_uy_k_dt_new), c_loc(d_E_k_dt_new), &
       c_loc(d_Bx_k_dt_new), c_loc(d_By_k_dt_new)]), 'derive2com')
end subroutine derive2com

subroutine ab2com(Nx, Ny, Nh, pi, time, rho_k, rho_ux_k, rho_uy_k, E_k, Bx_k, By_k, &
                  rho_k_new, rho_ux_k_new, rho_uy_k_new, E_k_new, Bx_k_new, By_k_new, &
                  d_rho_k_dt_old, d_rho_k_dt_new, d_rho_ux_k_dt_old, d_rho_ux_k_dt_new, &
                  d_rho_uy_k_dt_old, d_rho_uy_k_dt_new, d_E_k_dt_old, d_E_k_dt_new, &
                  d_Bx_k_dt_old, d_Bx_k_dt_new, d_By_k_dt_old, d_By_k_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nh
  real(kind=8) :: pi, time, dt, Lx, Ly, mu, ms, CS, mu_0, eta, spheat, kf, A
  complex(kind=8), target :: rho_k(Nh, Ny), rho_ux_k(Nh, Ny), rho_uy_k(Nh, Ny), E_k(Nh, Ny), Bx_k(Nh, Ny), By_k(Nh, Ny)
  complex(kind=8), target :: rho_k_new(Nh, Ny), rho_ux_k_new(Nh, Ny), rho_uy_k_new(Nh, Ny), E_k_new(Nh, Ny)
  complex(kind=8), target :: Bx_k_new(Nh, Ny), By_k_new(Nh, Ny)
  complex(kind=8), target :: d_rho_k_dt_old(Nh, Ny), d_rho_ux_k_dt_old(Nh, Ny), d_rho_uy_k_dt_old(Nh, Ny)
  complex(kind=8), target :: d_E_k_dt_old(Nh, Ny), d_Bx_k_dt_old(Nh, Ny), d_By_k_dt_old(Nh, Ny)
  complex(kind=8), target :: d_rho_k_dt_new(Nh, Ny), d_rho_ux_k_dt_new(Nh, Ny), d_rho_uy_k_dt_new(Nh, Ny)
  complex(kind=8), target :: d_E_k_dt_new(Nh, Ny), d_Bx_k_dt_new(Nh, Ny), d_By_k_dt_new(Nh, Ny)
  common/comm/Lx, Ly, spheat, mu, ms, CS, mu_0, eta, dt, kf, A
  call tara_check(tara_ab2com(Nx, Ny, Nh, dt, &
      [c_loc(rho_k), c_loc(rho_ux_k), c_loc(rho_uy_k), c_loc(E_k), c_loc(Bx_k), c_loc(By_k)], &
      [c_loc(rho_k_new), c_loc(rho_ux_k_new), c_loc(rho_uy_k_new), c_loc(E_k_new), c_loc(Bx_k_new), &
       c_loc(By_k_new)], &
      [c_loc(d_rho_k_dt_old), c_loc(d_rho_ux_k_dt_old), c_loc(d_rho_uy_k_dt_old), c_loc(d_E_k_dt_old), &
       c_loc(d_Bx_k_dt_old), c_loc(d_By_k_dt_old)], &
      [c_loc(d_rho_k_dt_new), c_loc(d_rho_ux_k_dt_new), c_loc(d_rho_uy_k_dt_new), c_loc(d_E_k_dt_new), &
       c_loc(d_Bx_k_dt_new), c_loc(d_By_k_dt_new)]), 'ab2com')
end subroutine ab2com
