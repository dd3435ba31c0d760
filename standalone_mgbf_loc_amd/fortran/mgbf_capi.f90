!-----------------------------------------------------------------------
! ISO_C_BINDING interface to libmgbf (include/mgbf.h) and to the few HIP
! runtime calls the Fortran host layer needs for device buffers.
!-----------------------------------------------------------------------
module mgbf_capi
use iso_c_binding
implicit none

integer(c_int), parameter :: MGBF_OK = 0
integer(c_int), parameter :: MGBF_OP_WAWT = 0, MGBF_OP_WT = 1, MGBF_OP_W = 2, &
                             MGBF_OP_A = 3, MGBF_OP_ASQRT = 4,              &
                             MGBF_OP_BSQRT = 5, MGBF_OP_BSQRTT = 6,         &
                             MGBF_OP_A2F = 7, MGBF_OP_F2A = 8, &
                             MGBF_OP_A2F_ALL = 9, MGBF_OP_F2A_ALL = 10
integer(c_int), parameter :: MGBF_NTIMERS = 6
integer(c_int), parameter :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2

! Same member order as struct mgbf_config (every &PARAMETERS_MGBETA key)
type, bind(C) :: mgbf_config_t
  real(c_double) :: mg_ampl01, mg_ampl02, mg_ampl03
  real(c_double) :: mg_weig1, mg_weig2, mg_weig3, mg_weig4, mg_weig5
  integer(c_int) :: hx, hy, hz, p
  integer(c_int) :: mgbf_line, mgbf_proc, n_ens
  integer(c_int) :: lm_a, lm, km2, km3
  integer(c_int) :: l_loc, l_loc_vertical, l_filt_g1, l_lin_vertical,       &
                    l_lin_horizontal, l_quad_horizontal, l_new_map,          &
                    l_vertical_filter, ldelta, lquart, lhelm
  integer(c_int) :: gm_max
  integer(c_int) :: nm0, mm0
  integer(c_int) :: nxPE, nyPE
  integer(c_int) :: im_filt, jm_filt
end type mgbf_config_t

type, bind(C) :: mgbf_dims_t
  integer(c_int) :: gm, nx, ny, km, km_all, km3_all, ntaps2, ntaps1
  integer(c_long_long) :: field_size, mg_size
  integer(c_int) :: na, ma, km_a_all
  integer(c_long_long) :: anal_size
  integer(c_int) :: ntaps3
end type mgbf_dims_t

interface
  integer(c_int) function mgbf_config_default(cfg) bind(C, name="mgbf_config_default")
    import :: c_int, mgbf_config_t
    type(mgbf_config_t), intent(out) :: cfg
  end function
  integer(c_int) function mgbf_derive(cfg, dims) bind(C, name="mgbf_derive")
    import :: c_int, mgbf_config_t, mgbf_dims_t
    type(mgbf_config_t), intent(in) :: cfg
    type(mgbf_dims_t), intent(out) :: dims
  end function
  integer(c_int) function mgbf_init(cfg, device, ctx) bind(C, name="mgbf_init")
    import :: c_int, c_ptr, mgbf_config_t
    type(mgbf_config_t), intent(in) :: cfg
    integer(c_int), value :: device
    type(c_ptr), intent(out) :: ctx
  end function
  integer(c_int) function mgbf_finalize(ctx) bind(C, name="mgbf_finalize")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  integer(c_int) function mgbf_get_dims(ctx, dims) bind(C, name="mgbf_get_dims")
    import :: c_int, c_ptr, mgbf_dims_t
    type(c_ptr), value :: ctx
    type(mgbf_dims_t), intent(out) :: dims
  end function
  integer(c_int) function mgbf_set_weights(ctx, gen, field, value) bind(C, name="mgbf_set_weights")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    integer(c_int), value :: gen
    type(c_ptr), value :: field
    real(c_double), value :: value
  end function
  integer(c_int) function mgbf_apply(ctx, op, in, out, stream) bind(C, name="mgbf_apply")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: op
    type(c_ptr), value :: in, out, stream
  end function
  integer(c_int) function mgbf_timers_enable(ctx, enable) bind(C, name="mgbf_timers_enable")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int), value :: enable
  end function
  integer(c_int) function mgbf_timers_read(ctx, ms, launches, bytes) bind(C, name="mgbf_timers_read")
    import :: c_int, c_ptr, c_double, c_long_long
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: ms(*)
    integer(c_long_long), intent(out) :: launches(*)
    real(c_double), intent(out) :: bytes(*)
  end function
  type(c_ptr) function mgbf_last_error_c() bind(C, name="mgbf_last_error")
    import :: c_ptr
  end function
  ! HIP runtime (device buffers of the host layer)
  integer(c_int) function hipMalloc(ptr, bytes) bind(C, name="hipMalloc")
    import :: c_int, c_ptr, c_size_t
    type(c_ptr), intent(out) :: ptr
    integer(c_size_t), value :: bytes
  end function
  integer(c_int) function hipFree(ptr) bind(C, name="hipFree")
    import :: c_int, c_ptr
    type(c_ptr), value :: ptr
  end function
  integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
    import :: c_int, c_ptr, c_size_t
    type(c_ptr), value :: dst, src
    integer(c_size_t), value :: bytes
    integer(c_int), value :: kind
  end function
  integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
    import :: c_int
  end function
end interface

contains

! Message of the last libmgbf failure as a Fortran string.
function mgbf_last_error() result(msg)
  character(len=:), allocatable :: msg
  type(c_ptr) :: p
  character(kind=c_char), pointer :: s(:)
  integer :: n
  interface
    integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: str
    end function
  end interface
  p = mgbf_last_error_c()
  n = int(c_strlen(p))
  call c_f_pointer(p, s, [n])
  allocate(character(len=n) :: msg)
  msg = transfer(s(1:n), msg)
end function mgbf_last_error

! Stops with the library's message when a call failed, as the reference
! stops on configuration errors.
subroutine mgbf_check(status, where)
  integer(c_int), intent(in) :: status
  character(len=*), intent(in) :: where
  if (status /= MGBF_OK) then
    write(*, '(a,": ",a)') where, mgbf_last_error()
    stop 'mgbf: call failed'
  endif
end subroutine mgbf_check

end module mgbf_capi
