! Test fixture: the three mod_parfile values the solver shim reads.  In psdm_run the
! reference's modules.f90 provides this module; the GPU test caller has no reference.
module mod_parfile
  implicit none
  integer :: dPML = 20, fs = 1, taper_option = 0
end module mod_parfile
