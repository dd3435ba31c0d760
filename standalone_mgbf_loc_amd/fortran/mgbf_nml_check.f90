!-----------------------------------------------------------------------
! Reads a namelist with init_mg_parameter and prints the sizes libmgbf
! derives from it (no device use): the check that the Fortran record of
! &PARAMETERS_MGBETA and the C struct agree.
!   mgbf_nml_check namelist
!-----------------------------------------------------------------------
program mgbf_nml_check
use iso_c_binding
use mgbf_capi
use mg_parameter, only: init_mg_parameter
use mg_entrymod, only: mg_config
implicit none
character(len=512) :: nml
type(mgbf_dims_t) :: d
integer(c_int) :: st
call get_command_argument(1, nml)
call init_mg_parameter(trim(nml))
st = mgbf_derive(mg_config(), d)
if (st /= MGBF_OK) then
  write(*, '(a,i0,a,a)') 'status ', st, ' ', mgbf_last_error()
  stop 1
endif
write(*, '(a,9(1x,i0))') 'dims', d%gm, d%nx, d%ny, d%km, d%km_all, d%km3_all, d%ntaps2, d%ntaps1, d%mg_size
end program mgbf_nml_check
