! Test caller of the Fortran SU reader binding: the driver's read of one MCS shot
! (calculate_PSDM_RTM.f90:212-214: open_shot_mcs + get_shot_data) as one call of
! psdm_get_shot_data (psdm_rtm_fortran.f90 -> libpsdm_rtm.so).
! Usage: su_caller folder file NumRec nt_max nt endianness delay dt out.bin
!   out.bin: float32 raw_data(nt,NumRec), then delay_data(nt,NumRec)
program su_caller
  use mod_parfile
  implicit none
  character(len=500) :: a, fout
  real, allocatable :: raw(:,:), delay(:,:)
  integer :: u
  call get_command_argument(1, folder_input_mcs)
  allocate(su_file_mcs(1))
  call get_command_argument(2, su_file_mcs(1))
  call get_command_argument(3, a); read(a, *) NumRec_MCS
  call get_command_argument(4, a); read(a, *) nt_mcs_max
  call get_command_argument(5, a); read(a, *) nt_mcs
  call get_command_argument(6, a); read(a, *) endianness_data
  call get_command_argument(7, a); read(a, *) delay_time_data
  call get_command_argument(8, a); read(a, *) dt_mcs
  call get_command_argument(9, fout)
  allocate(raw(nt_mcs, NumRec_MCS), delay(nt_mcs, NumRec_MCS))
  raw = -1.; delay = -1.
  call psdm_get_shot_data(1, raw, delay)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) raw, delay
  close(u)
end program
