! Test fixture: the shot window of dynamic_window (dynamic_window.f90:1-52) for one
! source, so that the GPU test caller links without the reference -- the source and
! receiver columns widened by 1 + floor(dd_space/dd) cells, clipped to 1..nx.  In psdm_run
! the reference's own dynamic_window.f90 provides it.
subroutine dynamic_window(rank, NumSou, NumRec, nxSou, nxRec, dd, nx, dd_space, nx_w, x_i, x_f)
  implicit none
  integer :: rank, NumSou, NumRec, nx, nx_w, x_i, x_f, ddx, dmin, dmax
  integer :: nxSou(NumSou), nxRec(NumRec, NumSou)
  real :: dd, dd_space
  ddx = 1 + floor(dd_space / dd)
  dmin = min(nxSou(1), minval(nxRec(:, 1)))
  dmax = max(nxSou(1), maxval(nxRec(:, 1)))
  x_i = max(dmin - ddx + 1, 1)
  x_f = min(dmax + ddx, nx)
  nx_w = x_f - x_i + 1
end subroutine
