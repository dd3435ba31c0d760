!-----------------------------------------------------------------------
! Host layer of the MGBF localization operator: the reference's module
! call surface (mg_parameter, mg_intstate, mg_entrymod, mg_transfer,
! mg_filtering) over the C ABI of libmgbf. One process drives one GPU and
! holds the whole domain: WORKA is the global analysis field
! WORKA(km_a_all, nm0, mm0) (the reference holds one PE tile of it).
!-----------------------------------------------------------------------

!=======================================================================
! Kind parameters (the reference's kinds module: r_kind = 8-byte real)
!=======================================================================
module mgbf_kinds
use iso_c_binding, only: c_double, c_int
implicit none
integer, parameter :: r_kind = c_double
integer, parameter :: i_kind = c_int
end module mgbf_kinds

!=======================================================================
! Namelist &PARAMETERS_MGBETA and derived sizes
! (src.mgbf/mg_parameter.f90:27-212, derived sizes :293-575)
!=======================================================================
module mg_parameter
use mgbf_kinds, only: r_kind, i_kind
implicit none
real(r_kind) :: mg_ampl01, mg_ampl02, mg_ampl03
real(r_kind) :: mg_weig1, mg_weig2, mg_weig3, mg_weig4, mg_weig5
integer(i_kind) :: mgbf_proc
logical :: mgbf_line
integer(i_kind) :: nxPE, nyPE, im_filt, jm_filt
logical :: lquart, lhelm
integer(i_kind) :: gm, gm_max
integer(i_kind) :: nm0, mm0
integer(i_kind) :: hx, hy, hz, p
integer(i_kind) :: lm_a, lm, km2, km3, n_ens
logical :: l_loc, l_loc_vertical, l_filt_g1, l_lin_vertical, l_lin_horizontal, &
           l_quad_horizontal, l_new_map, l_vertical_filter, ldelta
! derived
integer(i_kind) :: nm, mm, im, jm, im00, jm00
integer(i_kind) :: km, km_a, km_all, km_a_all, km2_all, km3_all

contains

subroutine init_mg_parameter(nml_file)
  character(len=*), intent(in), optional :: nml_file
  integer :: u
  namelist /parameters_mgbeta/ mg_ampl01, mg_ampl02, mg_ampl03,             &
                               mg_weig1, mg_weig2, mg_weig3, mg_weig4,      &
                               mg_weig5,                                    &
                               hx, hy, hz, p,                               &
                               mgbf_line, mgbf_proc, n_ens,                 &
                               lm_a, lm,                                    &
                               km2, km3,                                    &
                               l_loc, l_loc_vertical, l_filt_g1,            &
                               l_lin_vertical, l_lin_horizontal,            &
                               l_quad_horizontal, l_new_map,                &
                               l_vertical_filter,                           &
                               ldelta, lquart, lhelm,                       &
                               gm_max,                                      &
                               nm0, mm0,                                    &
                               nxPE, nyPE, im_filt, jm_filt
  ! keys the namelist may omit
  mg_ampl03 = 1.; mg_weig5 = 0.
  mgbf_line = .false.; l_loc = .false.; l_loc_vertical = .false.
  l_lin_vertical = .false.; l_lin_horizontal = .false.; l_quad_horizontal = .false.
  l_new_map = .true.; l_vertical_filter = .true.; ldelta = .false.
  lquart = .false.; lhelm = .false.; l_filt_g1 = .true.; km2 = 0
  if (present(nml_file)) then
    open(newunit=u, file=nml_file, status='old', action='read')
  else
    open(newunit=u, file='mgbeta.nml', status='old', action='read')
  endif
  read(u, nml=parameters_mgbeta)
  close(u)
  km_a = km2 + lm_a*km3
  km   = km2 + lm*km3
  km_a_all = km_a*n_ens
  km_all   = km*n_ens
  km2_all  = km2*n_ens
  km3_all  = km3*n_ens
  nm = nm0/nxPE
  mm = mm0/nyPE
  im = im_filt
  jm = jm_filt
  im00 = nxPE*im
  jm00 = nyPE*jm
end subroutine init_mg_parameter

end module mg_parameter

!=======================================================================
! State: the analysis field WORKA (src.mgbf/mg_intstate.f90:99,210) and
! the device-side filter field VALL
!=======================================================================
module mg_intstate
use iso_c_binding
use mgbf_kinds, only: r_kind
implicit none
real(r_kind), allocatable, target :: WORKA(:,:,:)
type(c_ptr) :: mgbf_ctx = c_null_ptr      ! libmgbf context
type(c_ptr) :: d_vall = c_null_ptr        ! device generation-1 field (FP32)
type(c_ptr) :: d_worka = c_null_ptr       ! device analysis field (FP32)
real(c_float), allocatable, target :: stage4(:,:,:)   ! FP32 host staging of WORKA
integer(c_size_t) :: vall_bytes = 0, worka_bytes = 0
end module mg_intstate

!=======================================================================
! mg_initialize / mg_finalize (src.mgbf/mg_entrymod.f90:36-273)
!=======================================================================
module mg_entrymod
use iso_c_binding
use mgbf_capi
use mg_parameter
use mg_intstate
implicit none
private
public :: mg_initialize, mg_finalize, mg_config

contains

! The C configuration record of the namelist values.
function mg_config() result(cfg)
  type(mgbf_config_t) :: cfg
  cfg%mg_ampl01 = mg_ampl01; cfg%mg_ampl02 = mg_ampl02; cfg%mg_ampl03 = mg_ampl03
  cfg%mg_weig1 = mg_weig1; cfg%mg_weig2 = mg_weig2; cfg%mg_weig3 = mg_weig3
  cfg%mg_weig4 = mg_weig4; cfg%mg_weig5 = mg_weig5
  cfg%hx = hx; cfg%hy = hy; cfg%hz = hz; cfg%p = p
  cfg%mgbf_line = merge(1, 0, mgbf_line); cfg%mgbf_proc = mgbf_proc; cfg%n_ens = n_ens
  cfg%lm_a = lm_a; cfg%lm = lm; cfg%km2 = km2; cfg%km3 = km3
  cfg%l_loc = merge(1, 0, l_loc); cfg%l_loc_vertical = merge(1, 0, l_loc_vertical)
  cfg%l_filt_g1 = merge(1, 0, l_filt_g1); cfg%l_lin_vertical = merge(1, 0, l_lin_vertical)
  cfg%l_lin_horizontal = merge(1, 0, l_lin_horizontal)
  cfg%l_quad_horizontal = merge(1, 0, l_quad_horizontal)
  cfg%l_new_map = merge(1, 0, l_new_map); cfg%l_vertical_filter = merge(1, 0, l_vertical_filter)
  cfg%ldelta = merge(1, 0, ldelta); cfg%lquart = merge(1, 0, lquart); cfg%lhelm = merge(1, 0, lhelm)
  cfg%gm_max = gm_max; cfg%nm0 = nm0; cfg%mm0 = mm0
  cfg%nxPE = nxPE; cfg%nyPE = nyPE; cfg%im_filt = im_filt; cfg%jm_filt = jm_filt
end function mg_config

subroutine mg_initialize(nml_file, device)
  character(len=*), intent(in), optional :: nml_file
  integer, intent(in), optional :: device
  type(mgbf_config_t) :: cfg
  type(mgbf_dims_t) :: dims
  integer(c_int) :: dev
  integer :: kd, id, jd
  call init_mg_parameter(nml_file)
  cfg = mg_config()
  dev = 0
  if (present(device)) dev = device
  call mgbf_check(mgbf_init(cfg, dev, mgbf_ctx), 'mg_initialize')
  call mgbf_check(mgbf_get_dims(mgbf_ctx, dims), 'mg_initialize')
  gm = dims%gm
  allocate(WORKA(km_a_all, nm0, mm0))
  allocate(stage4(km_a_all, nm0, mm0))
  WORKA = 0.
  vall_bytes = int(dims%field_size, c_size_t)*4_c_size_t
  worka_bytes = int(dims%anal_size, c_size_t)*4_c_size_t
  if (hipMalloc(d_vall, vall_bytes) /= 0) stop 'mg_initialize: hipMalloc failed'
  if (hipMalloc(d_worka, worka_bytes) /= 0) stop 'mg_initialize: hipMalloc failed'
  if (ldelta) then
    ! unit impulse of input_spec1_2d(..,'md') on the PE (nxPE/2, nyPE/2)
    ! into stack level 3*lm_a+lm_a/2 (src.mgbf/mg_entrymod.f90:198-213)
    kd = 3*lm_a + lm_a/2
    if (kd > km_a_all) stop 'mg_initialize: ldelta needs km3 >= 4'
    id = (nxPE/2 - 1)*nm + nm/2
    jd = (nyPE/2 - 1)*mm + mm/2
    WORKA(kd, id, jd) = 1.
  endif
end subroutine mg_initialize

subroutine mg_finalize
  integer(c_int) :: st
  if (c_associated(d_vall)) st = hipFree(d_vall)
  d_vall = c_null_ptr
  if (c_associated(d_worka)) st = hipFree(d_worka)
  d_worka = c_null_ptr
  if (c_associated(mgbf_ctx)) st = mgbf_finalize(mgbf_ctx)
  mgbf_ctx = c_null_ptr
  if (allocated(WORKA)) deallocate(WORKA)
  if (allocated(stage4)) deallocate(stage4)
end subroutine mg_finalize

end module mg_entrymod

!=======================================================================
! Analysis <-> filter grid (src.mgbf/mg_transfer.f90:58-203): WORKA goes
! to the device through an FP32 staging array, and the library maps it to
! the generation-1 field (anal_to_filt_all2 / anal_to_filt_all: the
! transpose of the horizontal and vertical interpolation) and back
! (filt_to_anal_all2 / filt_to_anal_all). The _all2 routines map the levels
! linearly, the _all routines linearly (l_lin_vertical) or with the lwq
! weights. With nm = im, mm = jm and lm_a = lm every map is a copy.
!=======================================================================
module mg_transfer
use iso_c_binding
use mgbf_capi
use mg_parameter
use mg_intstate
implicit none
private
public :: anal_to_filt_all, filt_to_anal_all, anal_to_filt_all2, filt_to_anal_all2

contains

subroutine to_filter(op)
  integer(c_int), intent(in) :: op
  stage4 = real(WORKA, c_float)
  if (hipMemcpy(d_worka, c_loc(stage4), worka_bytes, hipMemcpyHostToDevice) /= 0) &
    stop 'anal_to_filt: hipMemcpy failed'
  call mgbf_check(mgbf_apply(mgbf_ctx, op, d_worka, d_vall, c_null_ptr), 'anal_to_filt')
end subroutine to_filter

subroutine to_analysis(op)
  integer(c_int), intent(in) :: op
  call mgbf_check(mgbf_apply(mgbf_ctx, op, d_vall, d_worka, c_null_ptr), 'filt_to_anal')
  if (hipMemcpy(c_loc(stage4), d_worka, worka_bytes, hipMemcpyDeviceToHost) /= 0) &
    stop 'filt_to_anal: hipMemcpy failed'
  WORKA = real(stage4, r_kind)
end subroutine to_analysis

subroutine anal_to_filt_all2
  call to_filter(MGBF_OP_A2F)
end subroutine anal_to_filt_all2

subroutine filt_to_anal_all2
  call to_analysis(MGBF_OP_F2A)
end subroutine filt_to_anal_all2

subroutine anal_to_filt_all
  call to_filter(MGBF_OP_A2F_ALL)
end subroutine anal_to_filt_all

subroutine filt_to_anal_all
  call to_analysis(MGBF_OP_F2A_ALL)
end subroutine filt_to_anal_all

end module mg_transfer

!=======================================================================
! filtering_procedure (src.mgbf/mg_filtering.f90:67-120): W A W^T of the
! procedure the context was built for (1: rad2_bckg, 2: lin2_bckg,
! 3: fast_bckg, 4: rad3, 5: lin3, 6: rad2_ens, 7: lin2_ens, 8: fast_ens)
!=======================================================================
module mg_filtering
use iso_c_binding
use mgbf_capi
use mg_parameter, only: mgbf_proc
use mg_intstate
implicit none
private
public :: filtering_procedure

contains

subroutine filtering_procedure(mg_filt)
  integer, intent(in) :: mg_filt
  if (mg_filt /= mgbf_proc) stop 'filtering_procedure: procedure differs from mgbf_proc of mg_initialize'
  call mgbf_check(mgbf_apply(mgbf_ctx, MGBF_OP_WAWT, d_vall, d_vall, c_null_ptr), &
                  'filtering_procedure')
  if (hipDeviceSynchronize() /= 0) stop 'filtering_procedure: device failure'
end subroutine filtering_procedure

end module mg_filtering
