! Test fixture: the mod_parfile values the shim reads (the solvers' dPML, fs, taper_option,
! T_taper;
! the SU reader's file, byte order and sizes).  In psdm_run the reference's modules.f90
! provides this module; the GPU test callers have no reference.
module mod_parfile
  implicit none
  integer :: dPML = 20, fs = 1, taper_option = 0
  real :: T_taper = 0.05
  character(len=500) :: folder_input_mcs = ' '
  character(len=100), allocatable :: su_file_mcs(:)
  integer :: endianness_data = 1, NumRec_MCS = 0, nt_mcs = 0, nt_mcs_max = 0
  real :: delay_time_data = 0., dt_mcs = 1.
end module mod_parfile
