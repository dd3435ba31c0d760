!-----------------------------------------------------------------------
! Standalone driver with the call sequence of the reference's RBETA_TEST
! (src.mgbf/RBETA_TEST.f90:34-126): mg_initialize, anal_to_filt_all2,
! filtering_procedure(mgbf_proc), filt_to_anal_all2, mg_finalize.
!
!   mgbf_driver [namelist [input.bin [output.bin]]]
!
! input.bin / output.bin: raw FP64 WORKA(km_a_all, nm0, mm0) in Fortran
! order. Without an input file the driver filters the namelist's field
! (zero, or the ldelta impulse).
!-----------------------------------------------------------------------
program mgbf_driver
use iso_c_binding
use mg_parameter, only: mgbf_proc, l_new_map
use mg_intstate, only: WORKA
use mg_entrymod, only: mg_initialize, mg_finalize
use mg_transfer, only: anal_to_filt_all2, filt_to_anal_all2, anal_to_filt_all, filt_to_anal_all
use mg_filtering, only: filtering_procedure
implicit none
character(len=512) :: nml, fin, fout
integer :: u, nargs
logical :: have_in

nargs = command_argument_count()
nml = 'mgbeta.nml'; fin = ''; fout = 'output.bin'
if (nargs >= 1) call get_command_argument(1, nml)
if (nargs >= 2) call get_command_argument(2, fin)
if (nargs >= 3) call get_command_argument(3, fout)

call mg_initialize(trim(nml))
have_in = len_trim(fin) > 0
if (have_in) then
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) WORKA
  close(u)
endif

if (l_new_map) then
  call anal_to_filt_all2
else
  call anal_to_filt_all
endif
call filtering_procedure(mgbf_proc)
if (l_new_map) then
  call filt_to_anal_all2
else
  call filt_to_anal_all
endif

open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
write(u) WORKA
close(u)
write(*, '(a,i0,a,es12.5)') 'mgbf_driver: mgbf_proc=', mgbf_proc, ' max|WORKA|=', maxval(abs(WORKA))
call mg_finalize
end program mgbf_driver
