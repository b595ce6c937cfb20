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

    ! Device-resident mode (include/tara_hip.h): the state stays on the GPU
    ! between steps; handles are opaque c_ptr values.
    integer(c_int) function tara_csh_create(nx, ny, nu, dt, scheme, ctx) bind(C, name="tara_csh_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, scheme
      real(c_double), value :: nu, dt
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function tara_csh_upload(ctx, omegak) bind(C, name="tara_csh_upload")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: omegak(*)
    end function

    integer(c_int) function tara_csh_download(ctx, omegak) bind(C, name="tara_csh_download")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: omegak(*)
    end function

    integer(c_int) function tara_csh_advance(ctx, nsteps) bind(C, name="tara_csh_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function tara_csh_sync(ctx) bind(C, name="tara_csh_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function tara_csh_destroy(ctx) bind(C, name="tara_csh_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function tara_csh_ic_shear(ctx, lx_in, ly_in, w0, m0, d_in) bind(C, name="tara_csh_ic_shear")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: lx_in, ly_in, w0, m0, d_in
    end function

    integer(c_int) function tara_csh_fields(ctx, omega, ux, uy) bind(C, name="tara_csh_fields")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: omega(*), ux(*), uy(*)
    end function

    integer(c_int) function tara_csh_advance_tracers(ctx, nsteps, dt, lx, ly, periodic, np, xp, yp) &
        bind(C, name="tara_csh_advance_tracers")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, periodic, np
      real(c_double), value :: dt, lx, ly
      real(c_double), intent(inout) :: xp(*), yp(*)
    end function

    integer(c_int) function tara_mhd_create(nx, ny, nz, p, ctx) bind(C, name="tara_mhd_create")
      import :: c_int, c_ptr, tara_mhd_params
      integer(c_int), value :: nx, ny, nz
      type(tara_mhd_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function tara_mhd_upload(ctx, state) bind(C, name="tara_mhd_upload")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: state(*)
    end function

    integer(c_int) function tara_mhd_download(ctx, state) bind(C, name="tara_mhd_download")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: state(*)
    end function

    integer(c_int) function tara_mhd_advance(ctx, nsteps) bind(C, name="tara_mhd_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function tara_mhd_sync(ctx) bind(C, name="tara_mhd_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function tara_mhd_destroy(ctx) bind(C, name="tara_mhd_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function tara_mhd_ic_abc(ctx, rho0, u0, b0, p0, a, b, cc, kf) bind(C, name="tara_mhd_ic_abc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: rho0, u0, b0, p0, a, b, cc, kf
    end function

    integer(c_int) function tara_mhd_record_energy(ctx, nrec, every, rows) bind(C, name="tara_mhd_record_energy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrec, every
      real(c_double), intent(out) :: rows(*)
    end function

    ! checkpoints (include/tara_hip.h; host I/O)
    integer(c_int) function tara_ckpt_write(path, nx, ny, nz, nfields, step, time, fields) &
        bind(C, name="tara_ckpt_write")
      import :: c_int, c_long, c_double, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nx, ny, nz, nfields
      integer(c_long), value :: step
      real(c_double), value :: time
      type(c_ptr), intent(in) :: fields(*)
    end function

    integer(c_int) function tara_ckpt_read_header(path, nx, ny, nz, nfields, step, time) &
        bind(C, name="tara_ckpt_read_header")
      import :: c_int, c_long, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: nx, ny, nz, nfields
      integer(c_long), intent(out) :: step
      real(c_double), intent(out) :: time
    end function

    integer(c_int) function tara_ckpt_read(path, fields) bind(C, name="tara_ckpt_read")
      import :: c_int, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      type(c_ptr), intent(in) :: fields(*)
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

  ! File initial condition for a kept 2D main, in place of its commented reader
  ! (src/main2coh.f95:83, 149): reads a one-field checkpoint (written by
  ! tara_write_state2d, tara_ckpt_write or tara_amd.snapio) into omegak(Nh, Ny).
  subroutine tara_read_state2d(path, nh, ny, omegak, step, time)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nh, ny
    complex(c_double_complex), target, intent(inout) :: omegak(nh, ny)
    integer(c_long), intent(out), optional :: step
    real(c_double), intent(out), optional :: time
    character(kind=c_char, len=len_trim(path) + 1) :: cpath
    integer(c_int) :: fnx, fny, fnz, nf
    integer(c_long) :: st
    real(c_double) :: t
    type(c_ptr) :: p(1)
    cpath = trim(path) // c_null_char
    call tara_check(tara_ckpt_read_header(cpath, fnx, fny, fnz, nf, st, t), 'tara_read_state2d')
    if (fnx / 2 + 1 /= nh .or. fny /= ny .or. fnz /= 1 .or. nf /= 1) then
      write(0, '(a,a)') 'tara_read_state2d: field shape differs in ', trim(path)
      error stop 1
    end if
    p(1) = c_loc(omegak)
    call tara_check(tara_ckpt_read(cpath, p), 'tara_read_state2d')
    if (present(step)) step = st
    if (present(time)) time = t
  end subroutine

  ! The matching writer: omegak(Nx/2+1, Ny) with its step and time.
  subroutine tara_write_state2d(path, nx, ny, omegak, step, time)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny
    complex(c_double_complex), target, intent(in) :: omegak(nx / 2 + 1, ny)
    integer(c_long), intent(in) :: step
    real(c_double), intent(in) :: time
    character(kind=c_char, len=len_trim(path) + 1) :: cpath
    type(c_ptr) :: p(1)
    cpath = trim(path) // c_null_char
    p(1) = c_loc(omegak)
    call tara_check(tara_ckpt_write(cpath, int(nx, c_int), int(ny, c_int), 1_c_int, 1_c_int, step, time, p), &
                    'tara_write_state2d')
  end subroutine

end module tara_hip_mod
