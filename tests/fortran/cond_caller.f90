! Test caller of the Fortran conditioning binding: the driver's conditioning block
! (calculate_PSDM_RTM.f90:294-358: Data_target allocated and zeroed, then filter,
! resampling, cut_sf, taper / cut_multiple) as one call of psdm_condition_gather
! (psdm_rtm_fortran.f90 -> libpsdm_rtm.so).
! Usage: cond_caller in.bin out.bin
!   in.bin : int32 nt_data,NumRec,typef,nt_prop2,nt_taper,taper_option,N_taper,N_zeros,
!            cut_sf,cut_multiple; float32 dt_data,f_init,f_final,dt_prop;
!            int32 nt_sf(NumRec), nt_cut(NumRec); float32 Field_Data(nt_data,NumRec)
!   out.bin: float32 Data_target(nt_taper,NumRec)
program cond_caller
  use iso_c_binding, only: c_int
  implicit none
  integer :: nt_data, NumRec, typef, nt_prop2, nt_taper, taper_option, N_taper, N_zeros, cut_sf, cut_multiple
  integer :: u, ierr
  real :: dt_data, f_init, f_final, dt_prop
  integer(c_int), allocatable :: nt_sf(:), nt_cut(:)
  real, allocatable :: Field_Data(:,:), Data_target(:,:)
  character(len=500) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ierr)
  if (ierr /= 0) error stop 'cannot open input'
  read(u) nt_data, NumRec, typef, nt_prop2, nt_taper, taper_option, N_taper, N_zeros, cut_sf, cut_multiple
  read(u) dt_data, f_init, f_final, dt_prop
  allocate(nt_sf(NumRec), nt_cut(NumRec), Field_Data(nt_data,NumRec), Data_target(nt_taper,NumRec))
  read(u) nt_sf, nt_cut, Field_Data
  close(u)

  Data_target = 0
  call psdm_condition_gather(Field_Data, nt_data, dt_data, NumRec, typef, f_init, f_final, Data_target, &
                             nt_prop2, dt_prop, nt_taper, taper_option, N_taper, N_zeros, cut_sf, nt_sf, &
                             cut_multiple, nt_cut)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Data_target
  close(u)
end program
