! Test caller of the Fortran travel-time binding: the Kirchhoff driver's two calls
! (calculate_PSDM_kirchoff.f90:379-383) as psdm_window_tt_forward / _backward
! (psdm_rtm_fortran.f90 -> libpsdm_rtm.so).
! Usage: tt_caller in.bin out.bin
!   in.bin : int32 ny, nx, NumRec, nxSou, nySou, nxRec(NumRec), nyRec(NumRec); float32 dd,
!            dd_space, Vpmodel(ny, nx)
!   out.bin: int32 x_i, nx_w; float32 TT_s(ny, nx_w), TT_r(ny, nx_w, NumRec)
program tt_caller
  implicit none
  integer :: ny, nx, NumRec, u, nx_w, x_i, x_f, rank
  integer :: nxSou(1), nySou(1)
  integer, allocatable :: nxRec(:), nyRec(:)
  real :: dd, dd_space
  real, allocatable :: vel(:,:), Vp_w(:,:), TT_s(:,:), TT_r(:,:,:)
  character(len=500) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ny, nx, NumRec, nxSou(1), nySou(1)
  allocate(nxRec(NumRec), nyRec(NumRec), vel(ny, nx))
  read(u) nxRec, nyRec, dd, dd_space, vel
  close(u)
  rank = 0
  call dynamic_window(rank, 1, NumRec, nxSou, nxRec, dd, nx, dd_space, nx_w, x_i, x_f)
  allocate(Vp_w(ny, nx_w), TT_s(ny, nx_w), TT_r(ny, nx_w, NumRec))
  Vp_w(:, 1:nx_w) = vel(:, x_i:x_f)
  call psdm_window_tt_forward(rank, NumRec, nxSou, nySou, nxRec, nyRec, dd, ny, nx, nx_w, Vp_w, dd_space, TT_s)
  call psdm_window_tt_backward(rank, NumRec, nxSou, nySou, nxRec, nyRec, dd, ny, nx, nx_w, Vp_w, dd_space, TT_r)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) x_i, nx_w, TT_s, TT_r
  close(u)
end program
