! Test driver for libtara_dfftw.so: calls FFTW's Fortran entry points spelled
! as the reference's programs spell them (src/main2csh.f95:108-110,
! src/derive3com.f95:94-98, src/main1csh.f95:67-68), linked against the shim
! instead of FFTW.
!
! usage: drive_dfftw OUT
!   OUT: 2D r2c (32x16) of x + 2y + cos(x) sin(3y) pattern, its c2r, a 3D r2c
!   (16x8x8) and c2r, a 1D r2c/c2r (32), all as raw complex(8) / real(8).
program drive_dfftw
  implicit none
  integer, parameter :: FFTW_ESTIMATE = 64
  integer(kind=8) :: plan_f, plan_b, plan
  integer :: i, j, k, u, iret
  real(kind=8) :: r2(32,16), b2(32,16), r3(16,8,8), b3(16,8,8), r1(32), b1(32)
  complex(kind=8) :: c2(17,16), c3(9,8,8), c1(17)
  character(len=512) :: fout
  call get_command_argument(1, fout)
  do j = 1, 16
    do i = 1, 32
      r2(i,j) = dble(i) + 2.0d0*dble(j) + cos(dble(i))*sin(3.0d0*dble(j))
    end do
  end do
  do k = 1, 8
    do j = 1, 8
      do i = 1, 16
        r3(i,j,k) = sin(dble(i*j)) + 0.5d0*cos(dble(k + i))
      end do
    end do
  end do
  do i = 1, 32
    r1(i) = exp(-0.1d0*dble(i))
  end do
  call dfftw_init_threads(iret)
  call dfftw_plan_with_nthreads(4)
  call dfftw_plan_dft_r2c_2d_ (plan_f, 32, 16, r2, c2, FFTW_ESTIMATE)
  call dfftw_execute_ (plan_f)
  call dfftw_destroy_plan_ (plan_f)
  call dfftw_plan_dft_c2r_2d_ (plan_b, 32, 16, c2, b2, FFTW_ESTIMATE)
  call dfftw_execute_ (plan_b)
  call dfftw_destroy_plan_ (plan_b)
  call dfftw_plan_dft_r2c_3d_ (plan_f, 16, 8, 8, r3, c3, FFTW_ESTIMATE)
  call dfftw_execute_ (plan_f)
  call dfftw_destroy_plan_ (plan_f)
  call dfftw_plan_dft_c2r_3d_ (plan_b, 16, 8, 8, c3, b3, FFTW_ESTIMATE)
  call dfftw_execute_ (plan_b)
  call dfftw_destroy_plan_ (plan_b)
  call dfftw_plan_dft_r2c_1d(plan, 32, r1, c1, FFTW_ESTIMATE)
  call dfftw_execute_dft_r2c(plan, r1, c1)
  call dfftw_destroy_plan(plan)
  call dfftw_plan_dft_c2r_1d(plan, 32, c1, b1, FFTW_ESTIMATE)
  call dfftw_execute_dft_c2r(plan, c1, b1)
  call dfftw_destroy_plan(plan)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) r2, c2, b2, r3, c3, b3, r1, c1, b1
  close(u)
end program drive_dfftw
