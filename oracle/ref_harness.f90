! Test infrastructure only (oracle side).  Never linked into the product.
!
! C-callable driver loops around the reference's OWN routines (src/derive2csh.f95,
! src/ab2csh.f95, src/derive3com.f95, src/ab3com.f95) and the vendored FFTW, so the
! golden-vector generator and the CPU baseline can run the reference arithmetic
! without the reference's I/O-heavy main programs.  Each routine restates the
! driver logic it mirrors, cited per routine.  Arrays are Fortran column-major;
! complex(8) is interleaved (re, im).

module ref_harness
  use iso_c_binding
  implicit none
  include "fftw3.f"
  real(8), parameter :: pi = 3.14159265358979323846d0
contains

  ! Shear-layer initial vorticity and its r2c, as src/main2csh.f95:84-110.
  ! lx_in, ly_in are the input.ini values (the code multiplies them by pi,
  ! main2csh.f95:84-85) and d_in is multiplied by pi (main2csh.f95:91).
  subroutine ref_ic2csh(nx, ny, lx_in, ly_in, w0, m0, d_in, omega, omegak) bind(C, name="ref_ic2csh")
    integer(c_int), value :: nx, ny
    real(c_double), value :: lx_in, ly_in, w0, m0, d_in
    real(c_double) :: omega(nx, ny)
    complex(c_double_complex) :: omegak(nx/2+1, ny)
    real(8) :: lx, ly, dx, dy, d, x, y
    real(8), allocatable :: tmp(:,:)
    integer :: i, j
    integer(8) :: plan
    lx = lx_in*pi
    ly = ly_in*pi
    dx = lx/dfloat(nx)
    dy = ly/dfloat(ny)
    d = d_in*pi
    do i = 1, nx
      x = (i-1)*dx
      do j = 1, ny
        y = (j-1)*dy
        omega(i,j) = w0/dcosh((y+0.5d0*pi)/d)**2.0 - w0/dcosh((y-0.5d0*pi)/d)**2.0
        omega(i,j) = omega(i,j) - w0/dcosh((y+1.50d0*pi)/d)**2.0 + w0/dcosh((y-1.50d0*pi)/d)**2.0
        omega(i,j) = omega(i,j) + 0.10d0*dcos(m0*x)
      end do
    end do
    allocate(tmp(nx, ny))
    tmp = omega
    call dfftw_plan_dft_r2c_2d_(plan, nx, ny, tmp, omegak, FFTW_ESTIMATE)
    call dfftw_execute_(plan)
    call dfftw_destroy_plan_(plan)
    deallocate(tmp)
  end subroutine

  ! AB2 loop of src/main2csh.f95:114-126: derive2csh, ab2csh, history rotation.
  ! The history array dt_omegak_old is allocated uninitialised by the reference
  ! (main2csh.f95:64-65); here it is whatever the caller passes in hist (zero for
  ! the build's contract).  omegak is advanced in place by nsteps steps; hist holds
  ! the last RHS on return.
  subroutine ref_ab2csh_steps(nx, ny, nu, dt, nsteps, omegak, hist) bind(C, name="ref_ab2csh_steps")
    integer(c_int), value :: nx, ny, nsteps
    real(c_double), value :: nu, dt
    complex(c_double_complex) :: omegak(nx/2+1, ny), hist(nx/2+1, ny)
    complex(8), allocatable :: omegak_new(:,:), dt_new(:,:)
    integer :: t, nh
    real(8) :: time
    nh = nx/2 + 1
    allocate(omegak_new(nh, ny), dt_new(nh, ny))
    time = 0.0d0
    do t = 1, nsteps
      call derive2csh(nx, ny, nh, nu, time, omegak, omegak_new, hist, dt_new)
      call ab2csh(nx, ny, nh, nu, time, dt, omegak, omegak_new, hist, dt_new)
      hist = dt_new
      omegak = omegak_new
      time = time + dt
    end do
    deallocate(omegak_new, dt_new)
  end subroutine

  ! Classic RK4 over the reference right-hand side derive2csh.  The reference has
  ! no working RK4 for this family (src/rk42css.f95:15-48 states the intended
  ! 1/6(k1+2k2+2k3+k4) scheme with stages dt/2, dt/2, dt); this is that scheme.
  subroutine ref_rk4csh_steps(nx, ny, nu, dt, nsteps, omegak) bind(C, name="ref_rk4csh_steps")
    integer(c_int), value :: nx, ny, nsteps
    real(c_double), value :: nu, dt
    complex(c_double_complex) :: omegak(nx/2+1, ny)
    complex(8), allocatable :: w(:,:), k1(:,:), k2(:,:), k3(:,:), k4(:,:), dum(:,:)
    integer :: t, nh
    real(8) :: time
    nh = nx/2 + 1
    allocate(w(nh,ny), k1(nh,ny), k2(nh,ny), k3(nh,ny), k4(nh,ny), dum(nh,ny))
    time = 0.0d0
    do t = 1, nsteps
      call derive2csh(nx, ny, nh, nu, time, omegak, dum, dum, k1)
      w = omegak + (0.5d0*dt)*k1
      call derive2csh(nx, ny, nh, nu, time, w, dum, dum, k2)
      w = omegak + (0.5d0*dt)*k2
      call derive2csh(nx, ny, nh, nu, time, w, dum, dum, k3)
      w = omegak + dt*k3
      call derive2csh(nx, ny, nh, nu, time, w, dum, dum, k4)
      omegak = omegak + (dt/6.0d0)*(k1 + 2.0d0*k2 + 2.0d0*k3 + k4)
      time = time + dt
    end do
    deallocate(w, k1, k2, k3, k4, dum)
  end subroutine

  ! AB2 loop of src/main3com.f95:536-617: derive3com, ab3com, then the history
  ! and state rotation.  patched /= 0 calls derive3com_mueta, the reference's
  ! derive3com with its hard-wired 1/450 diffusion replaced by COMMON's mu and
  ! eta (oracle/Makefile); 0 calls the unmodified derive3com.  The physics
  ! parameters (Lx, Ly, Lz, CS, mu, eta, dt, thread_num) are read from COMMON
  ! /comm/, which the caller fills.  st(:,:,:,f) and hist(:,:,:,f) hold the
  ! fields in the argument order of derive3com (rho, rho ux, rho uy, rho uz, E,
  ! Bx, By, Bz); both are advanced in place (hist = the last RHS on return).
  subroutine ref_mhd_ab2_steps(nx, ny, nz, nsteps, patched, st, hist) bind(C, name="ref_mhd_ab2_steps")
    integer(c_int), value :: nx, ny, nz, nsteps, patched
    complex(c_double_complex) :: st(nx/2+1, ny, nz, 8), hist(nx/2+1, ny, nz, 8)
    integer :: iret, thread_num
    real(8) :: time_max, Lx, Ly, Lz, dx, dy, dz, spheat, mu, ms, CS, mu_0, eta, dt, kf
    common/comm/iret,thread_num,time_max,Lx,Ly,Lz,dx,dy,dz,spheat,mu,ms,CS,mu_0,eta,dt,kf
    complex(8), allocatable :: sn(:,:,:,:), fn(:,:,:,:)
    integer :: t, nh
    real(8) :: time
    nh = nx/2 + 1
    allocate(sn(nh, ny, nz, 8), fn(nh, ny, nz, 8))
    time = 0.0d0
    do t = 1, nsteps
      if (patched /= 0) then
        call derive3com_mueta(nx, ny, nz, nh, pi, time, st(:,:,:,1), st(:,:,:,2), st(:,:,:,3), st(:,:,:,4), &
                              st(:,:,:,5), st(:,:,:,6), st(:,:,:,7), st(:,:,:,8), &
                              hist(:,:,:,1), fn(:,:,:,1), hist(:,:,:,2), fn(:,:,:,2), hist(:,:,:,3), fn(:,:,:,3), &
                              hist(:,:,:,4), fn(:,:,:,4), hist(:,:,:,5), fn(:,:,:,5), &
                              hist(:,:,:,6), fn(:,:,:,6), hist(:,:,:,7), fn(:,:,:,7), hist(:,:,:,8), fn(:,:,:,8))
      else
        call derive3com(nx, ny, nz, nh, pi, time, st(:,:,:,1), st(:,:,:,2), st(:,:,:,3), st(:,:,:,4), &
                        st(:,:,:,5), st(:,:,:,6), st(:,:,:,7), st(:,:,:,8), &
                        hist(:,:,:,1), fn(:,:,:,1), hist(:,:,:,2), fn(:,:,:,2), hist(:,:,:,3), fn(:,:,:,3), &
                        hist(:,:,:,4), fn(:,:,:,4), hist(:,:,:,5), fn(:,:,:,5), &
                        hist(:,:,:,6), fn(:,:,:,6), hist(:,:,:,7), fn(:,:,:,7), hist(:,:,:,8), fn(:,:,:,8))
      end if
      call ab3com(nx, ny, nz, nh, pi, time, st(:,:,:,1), st(:,:,:,2), st(:,:,:,3), st(:,:,:,4), &
                  st(:,:,:,5), st(:,:,:,6), st(:,:,:,7), st(:,:,:,8), &
                  sn(:,:,:,1), sn(:,:,:,2), sn(:,:,:,3), sn(:,:,:,4), sn(:,:,:,5), sn(:,:,:,6), sn(:,:,:,7), &
                  sn(:,:,:,8), &
                  hist(:,:,:,1), fn(:,:,:,1), hist(:,:,:,2), fn(:,:,:,2), hist(:,:,:,3), fn(:,:,:,3), &
                  hist(:,:,:,4), fn(:,:,:,4), hist(:,:,:,5), fn(:,:,:,5), &
                  hist(:,:,:,6), fn(:,:,:,6), hist(:,:,:,7), fn(:,:,:,7), hist(:,:,:,8), fn(:,:,:,8))
      hist = fn
      st = sn
      time = time + dt
    end do
    deallocate(sn, fn)
  end subroutine

end module ref_harness
