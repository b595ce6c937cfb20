! Drop-in replacements for the reference's 3D compressible MHD routines
! (src/derive3com.f95:6, src/ab3com.f95:6) with their names, argument lists and
! COMMON /comm/ layout (src/main3com.f95:56).  The diffusion coefficients are
! 1/450 as the reference hard-wires them, or mu and eta from COMMON when
! tara_mhd_diffusion_from_common is set.

subroutine derive3com(Nx, Ny, Nz, Nh, pi, time, rho_k, rho_ux_k, rho_uy_k, rho_uz_k, E_k, Bx_k, By_k, Bz_k, &
                      d_rho_k_dt_old, d_rho_k_dt_new, d_rho_ux_k_dt_old, d_rho_ux_k_dt_new, &
                      d_rho_uy_k_dt_old, d_rho_uy_k_dt_new, d_rho_uz_k_dt_old, d_rho_uz_k_dt_new, &
                      d_E_k_dt_old, d_E_k_dt_new, d_Bx_k_dt_old, d_Bx_k_dt_new, &
                      d_By_k_dt_old, d_By_k_dt_new, d_Bz_k_dt_old, d_Bz_k_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nz, Nh, iret, thread_num
  real(kind=8) :: pi, time, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: rho_k, rho_ux_k, rho_uy_k, rho_uz_k, E_k, Bx_k, By_k, Bz_k
  complex(kind=8), dimension(Nh, Ny, Nz) :: d_rho_k_dt_old, d_rho_ux_k_dt_old, d_rho_uy_k_dt_old, &
      d_rho_uz_k_dt_old, d_E_k_dt_old, d_Bx_k_dt_old, d_By_k_dt_old, d_Bz_k_dt_old
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: d_rho_k_dt_new, d_rho_ux_k_dt_new, d_rho_uy_k_dt_new, &
      d_rho_uz_k_dt_new, d_E_k_dt_new, d_Bx_k_dt_new, d_By_k_dt_new, d_Bz_k_dt_new
  common/comm/iret, thread_num, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  type(tara_mhd_params) :: p
  if (tara_mhd_diffusion_from_common) then
    p = tara_mhd_params(Lx, Ly, Lz, mu, eta, CS, dt)
  else
    p = tara_mhd_params(Lx, Ly, Lz, 1.0d0/450.0d0, 1.0d0/450.0d0, CS, dt)
  end if
  call tara_check(tara_derive3com(Nx, Ny, Nz, Nh, p, &
      [c_loc(rho_k), c_loc(rho_ux_k), c_loc(rho_uy_k), c_loc(rho_uz_k), c_loc(E_k), c_loc(Bx_k), c_loc(By_k), &
       c_loc(Bz_k)], &
      [c_loc(d_rho_k_dt_new), c_loc(d_rho_ux_k_dt_new), c_loc(d_rho_uy_k_dt_new), c_loc(d_rho_uz_k_dt_new), &
       c_loc(d_E_k_dt_new), c_loc(d_Bx_k_dt_new), c_loc(d_By_k_dt_new), c_loc(d_Bz_k_dt_new)]), 'derive3com')
end subroutine derive3com

subroutine ab3com(Nx, Ny, Nz, Nh, pi, time, rho_k, rho_ux_k, rho_uy_k, rho_uz_k, E_k, Bx_k, By_k, Bz_k, &
                  rho_k_new, rho_ux_k_new, rho_uy_k_new, rho_uz_k_new, E_k_new, Bx_k_new, By_k_new, Bz_k_new, &
                  d_rho_k_dt_old, d_rho_k_dt_new, d_rho_ux_k_dt_old, d_rho_ux_k_dt_new, &
                  d_rho_uy_k_dt_old, d_rho_uy_k_dt_new, d_rho_uz_k_dt_old, d_rho_uz_k_dt_new, &
                  d_E_k_dt_old, d_E_k_dt_new, d_Bx_k_dt_old, d_Bx_k_dt_new, &
                  d_By_k_dt_old, d_By_k_dt_new, d_Bz_k_dt_old, d_Bz_k_dt_new)
  use tara_hip_mod
  implicit none
  integer(kind=4) :: Nx, Ny, Nz, Nh, iret, thread_num
  real(kind=8) :: pi, time, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: rho_k, rho_ux_k, rho_uy_k, rho_uz_k, E_k, Bx_k, By_k, Bz_k
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: rho_k_new, rho_ux_k_new, rho_uy_k_new, rho_uz_k_new, E_k_new, &
      Bx_k_new, By_k_new, Bz_k_new
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: d_rho_k_dt_old, d_rho_ux_k_dt_old, d_rho_uy_k_dt_old, &
      d_rho_uz_k_dt_old, d_E_k_dt_old, d_Bx_k_dt_old, d_By_k_dt_old, d_Bz_k_dt_old
  complex(kind=8), target, dimension(Nh, Ny, Nz) :: d_rho_k_dt_new, d_rho_ux_k_dt_new, d_rho_uy_k_dt_new, &
      d_rho_uz_k_dt_new, d_E_k_dt_new, d_Bx_k_dt_new, d_By_k_dt_new, d_Bz_k_dt_new
  common/comm/iret, thread_num, time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
  call tara_check(tara_ab3com(Nx, Ny, Nz, Nh, dt, &
      [c_loc(rho_k), c_loc(rho_ux_k), c_loc(rho_uy_k), c_loc(rho_uz_k), c_loc(E_k), c_loc(Bx_k), c_loc(By_k), &
       c_loc(Bz_k)], &
      [c_loc(rho_k_new), c_loc(rho_ux_k_new), c_loc(rho_uy_k_new), c_loc(rho_uz_k_new), c_loc(E_k_new), &
       c_loc(Bx_k_new), c_loc(By_k_new), c_loc(Bz_k_new)], &
      [c_loc(d_rho_k_dt_old), c_loc(d_rho_ux_k_dt_old), c_loc(d_rho_uy_k_dt_old), c_loc(d_rho_uz_k_dt_old), &
       c_loc(d_E_k_dt_old), c_loc(d_Bx_k_dt_old), c_loc(d_By_k_dt_old), c_loc(d_Bz_k_dt_old)], &
      [c_loc(d_rho_k_dt_new), c_loc(d_rho_ux_k_dt_new), c_loc(d_rho_uy_k_dt_new), c_loc(d_rho_uz_k_dt_new), &
       c_loc(d_E_k_dt_new), c_loc(d_Bx_k_dt_new), c_loc(d_By_k_dt_new), c_loc(d_Bz_k_dt_new)]), 'ab3com')
end subroutine ab3com
