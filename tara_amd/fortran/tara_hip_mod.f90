! ISO_C_BINDING interfaces to libtara_hip.so (include/tara_hip.h) and the
! error path shared by the drop-in wrappers: a nonzero status from the C ABI
! becomes `error stop` with the library's message.
module tara_hip_mod
  use iso_c_binding
  implicit none

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
