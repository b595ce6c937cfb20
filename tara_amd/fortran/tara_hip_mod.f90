! ISO_C_BINDING interfaces to libtara_hip.so (include/tara_hip.h) and the
! error path shared by the drop-in wrappers: a nonzero status from the C ABI
! becomes `error stop` with the library's message.
module tara_hip_mod
  use iso_c_binding
  implicit none

  ! Parameter blocks of include/tara_hip.h
  type, bind(C) :: tara_coh_params
    real(c_double) :: pi, lx, ly, mux, muy, cs, dt
  end type
  type, bind(C) :: tara_com_params
    real(c_double) :: pi, lx, ly, mu, eta, cs, kf, a, dt
  end type
  type, bind(C) :: tara_mhd_params
    real(c_double) :: lx, ly, lz, mu, eta, cs, dt
  end type

  ! lambda of the screened form: derive2css leaves it uninitialised; the
  ! reference driver sets 10 (src/main2css.f95:87).  Callers may change it.
  real(c_double) :: tara_css_lambda = 10.0d0
  ! derive3com hard-wires the diffusion coefficients to 1/450
  ! (src/derive3com.f95:1023-1037); set this to .true. to take mu and eta from
  ! COMMON /comm/ instead.
  logical :: tara_mhd_diffusion_from_common = .false.

  interface
    integer(c_int) function tara_derive2csh(nx, ny, nh, nu, omegak, dt_new) bind(C, name="tara_derive2csh")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: nu
      complex(c_double_complex), intent(in) :: omegak(*)
      complex(c_double_complex), intent(out) :: dt_new(*)
    end function

    integer(c_int) function tara_ab2csh(nx, ny, nh, dt, omegak, omegak_new, dt_old, dt_new) &
        bind(C, name="tara_ab2csh")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      complex(c_double_complex), intent(in) :: omegak(*), dt_old(*), dt_new(*)
      complex(c_double_complex), intent(out) :: omegak_new(*)
    end function

    integer(c_int) function tara_derive2css(nx, ny, nh, alpha, nu, lambda, omegak, dt_new) &
        bind(C, name="tara_derive2css")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: alpha, nu, lambda
      complex(c_double_complex), intent(in) :: omegak(*)
      complex(c_double_complex), intent(out) :: dt_new(*)
    end function

    integer(c_int) function tara_ab2css(nx, ny, nh, dt, omegak, omegak_new, dt_old, dt_new) &
        bind(C, name="tara_ab2css")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      complex(c_double_complex), intent(in) :: omegak(*), dt_old(*), dt_new(*)
      complex(c_double_complex), intent(out) :: omegak_new(*)
    end function

    integer(c_int) function tara_rk42css(nx, ny, nh, alpha, nu, lambda, dt, omegak, omegak_new) &
        bind(C, name="tara_rk42css")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: alpha, nu, lambda, dt
      complex(c_double_complex), intent(in) :: omegak(*)
      complex(c_double_complex), intent(out) :: omegak_new(*)
    end function

    integer(c_int) function tara_derive2coh(nx, ny, nh, p, state, rhs) bind(C, name="tara_derive2coh")
      import :: c_int, c_ptr, tara_coh_params
      integer(c_int), value :: nx, ny, nh
      type(tara_coh_params), intent(in) :: p
      type(c_ptr), intent(in) :: state(3), rhs(3)
    end function

    integer(c_int) function tara_ab2coh(nx, ny, nh, dt, state, state_new, f_old, f_new) bind(C, name="tara_ab2coh")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      type(c_ptr), intent(in) :: state(3), state_new(3), f_old(3), f_new(3)
    end function

    integer(c_int) function tara_euler2coh(nx, ny, nh, dt, state, state_new, slope) bind(C, name="tara_euler2coh")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      type(c_ptr), intent(in) :: state(3), state_new(3), slope(3)
    end function

    integer(c_int) function tara_rk4coh_field(nx, ny, nh, dt, field, field_new, slope) &
        bind(C, name="tara_rk4coh_field")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      complex(c_double_complex), intent(in) :: field(*), slope(*)
      complex(c_double_complex), intent(out) :: field_new(*)
    end function

    integer(c_int) function tara_pccoh_field(nx, ny, nh, dt, field, field_new, slope) &
        bind(C, name="tara_pccoh_field")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      complex(c_double_complex), intent(in) :: field(*), slope(*)
      complex(c_double_complex), intent(out) :: field_new(*)
    end function

    integer(c_int) function tara_derive2com(nx, ny, nh, p, state, rhs) bind(C, name="tara_derive2com")
      import :: c_int, c_ptr, tara_com_params
      integer(c_int), value :: nx, ny, nh
      type(tara_com_params), intent(in) :: p
      type(c_ptr), intent(in) :: state(6), rhs(6)
    end function

    integer(c_int) function tara_ab2com(nx, ny, nh, dt, state, state_new, f_old, f_new) bind(C, name="tara_ab2com")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nx, ny, nh
      real(c_double), value :: dt
      type(c_ptr), intent(in) :: state(6), state_new(6), f_old(6), f_new(6)
    end function

    integer(c_int) function tara_derive3com(nx, ny, nz, nh, p, state, rhs) bind(C, name="tara_derive3com")
      import :: c_int, c_ptr, tara_mhd_params
      integer(c_int), value :: nx, ny, nz, nh
      type(tara_mhd_params), intent(in) :: p
      type(c_ptr), intent(in) :: state(8), rhs(8)
    end function

    integer(c_int) function tara_ab3com(nx, ny, nz, nh, dt, state, state_new, f_old, f_new) &
        bind(C, name="tara_ab3com")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nx, ny, nz, nh
      real(c_double), value :: dt
      type(c_ptr), intent(in) :: state(8), state_new(8), f_old(8), f_new(8)
    end function

    type(c_ptr) function tara_last_error() bind(C, name="tara_last_error")
      import :: c_ptr
    end function
  end interface

contains

  subroutine tara_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (rc == 0) return
    call c_f_pointer(tara_last_error(), msg, [512])
    text = ' '
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write(0, '(a,a,i0,a,a)') trim(where), ': libtara_hip status ', rc, ': ', trim(text)
    error stop 1
  end subroutine

end module tara_hip_mod
