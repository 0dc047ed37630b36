! Fortran side of the drop-in: the four solver subroutines of PSDM's RTM path with the
! reference's own names and argument lists, forwarding to libpsdm_rtm.so through
! iso_c_binding (include/psdm_rtm.h).  Linking this file (with -lpsdm_rtm) instead of
! solver_ac_forward_ram.f90, solver_ac_backward_ram.f90, solver_ac_forward_disk.f90
! and solver_ac_backward_disk.f90 leaves their callers (dynamic_window.f90:90-92,
! 134-137, 177-179, 221-224) and calculate_PSDM_RTM.f90 unchanged.
!
! Module inputs are read exactly where the reference reads them: dPML and fs from
! mod_parfile (solver_ac_forward_ram.f90:8).  The forward snapshots stay in HBM until the
! backward call; with taper_option != 0 the driver's taper of Fields (calculate_PSDM_RTM.f90:
! 421-427, N_taper = max(1, int(T_taper/dt_prop)) at :287) is applied to them there at the
! end of the forward call, so the caller's Fields are never materialised.
! Errors keep the reference's fail-fast behaviour: a message, then error stop.

module psdm_rtm_iface
  use iso_c_binding
  implicit none
  type(c_ptr), save :: psdm_ctx = c_null_ptr

  ! psdm_rtm_cond of include/psdm_rtm.h (gather conditioning)
  type, bind(C) :: psdm_rtm_cond
    integer(c_int) :: nt_data
    real(c_float) :: dt_data
    integer(c_int) :: typef
    real(c_float) :: f_init, f_final
    integer(c_int) :: nt_prop2
    real(c_float) :: dt_prop
    integer(c_int) :: nt_taper, taper_option, N_taper, N_zeros
    type(c_ptr) :: nt_sf, nt_cut
  end type

  interface
    integer(c_int) function psdm_rtm_create(ctx, device) bind(C, name="psdm_rtm_create")
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function
    type(c_ptr) function psdm_rtm_last_error(ctx) bind(C, name="psdm_rtm_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function psdm_rtm_forward_ram(ctx, dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, &
        dmodel, dt, ny, nx, ctmp, store_snap, nt_snap, nt, Fields, host_fields, Rec_point) &
        bind(C, name="psdm_rtm_forward_ram")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: dPML, fs, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, host_fields
      integer(c_int) :: nxSou(*), nySou(*), nxRec(*), nyRec(*)
      real(c_float), value :: dmodel, dt
      real(c_float) :: S(*), ctmp(*), Fields(*), Rec_point(*)
    end function
    integer(c_int) function psdm_rtm_taper_snapshots(ctx, Fields, nt_prop2, N_taper) &
        bind(C, name="psdm_rtm_taper_snapshots")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: Fields(*)
      integer(c_int), value :: nt_prop2, N_taper
    end function
    integer(c_int) function psdm_rtm_backward_ram(ctx, dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, &
        dmodel, dt, ny, nx, ctmp, store_snap, nt_snap, nt, Fields, host_fields, kernel) &
        bind(C, name="psdm_rtm_backward_ram")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: dPML, fs, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, host_fields
      integer(c_int) :: nxSou(*), nySou(*), nxRec(*), nyRec(*)
      real(c_float), value :: dmodel, dt
      real(c_float) :: Sadj(*), ctmp(*), Fields(*), kernel(*)
    end function
    integer(c_int) function psdm_rtm_forward_disk(ctx, dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, &
        dmodel, dt, ny, nx, ctmp, shotid, store_snap, nt_snap, nt, Rec_point) bind(C, name="psdm_rtm_forward_disk")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: dPML, fs, NumSou, NumRec, ny, nx, shotid, store_snap, nt_snap, nt
      integer(c_int) :: nxSou(*), nySou(*), nxRec(*), nyRec(*)
      real(c_float), value :: dmodel, dt
      real(c_float) :: S(*), ctmp(*), Rec_point(*)
    end function
    integer(c_int) function psdm_rtm_backward_disk(ctx, dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, &
        dmodel, dt, ny, nx, ctmp, shotid, store_snap, nt_snap, nt, kernel) bind(C, name="psdm_rtm_backward_disk")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: dPML, fs, NumSou, NumRec, ny, nx, shotid, store_snap, nt_snap, nt
      integer(c_int) :: nxSou(*), nySou(*), nxRec(*), nyRec(*)
      real(c_float), value :: dmodel, dt
      real(c_float) :: Sadj(*), ctmp(*), kernel(*)
    end function
    integer(c_int) function psdm_rtm_su_read_host(ctx, path, offset, ntr, ns, nt, big_endian, shift, h_out) &
        bind(C, name="psdm_rtm_su_read_host")
      import :: c_ptr, c_int, c_float, c_char, c_long_long
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      integer(c_long_long), value :: offset
      integer(c_int), value :: ntr, ns, nt, big_endian, shift
      real(c_float) :: h_out(*)
    end function
    integer(c_int) function psdm_rtm_eikonal_host(ctx, h_vel, ny, nx, dx, dy, nsrc, src_x, src_z, h_tt) &
        bind(C, name="psdm_rtm_eikonal_host")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: h_vel(*), h_tt(*)
      integer(c_int), value :: ny, nx, nsrc
      real(c_float), value :: dx, dy
      integer(c_int) :: src_x(*), src_z(*)
    end function
    integer(c_int) function psdm_rtm_condition_host(ctx, cond, NumRec, Field_Data, Data_target) &
        bind(C, name="psdm_rtm_condition_host")
      import :: c_ptr, c_int, c_float, psdm_rtm_cond
      type(c_ptr), value :: ctx
      type(psdm_rtm_cond) :: cond
      integer(c_int), value :: NumRec
      real(c_float) :: Field_Data(*), Data_target(*)
    end function
  end interface

contains

  ! Context of this process (one rank per GPU: the launcher selects the device through
  ! HIP_VISIBLE_DEVICES, so the library always opens device 0).
  type(c_ptr) function psdm_context()
    integer(c_int) :: rc
    if (.not. c_associated(psdm_ctx)) then
      rc = psdm_rtm_create(psdm_ctx, 0_c_int)
      if (rc /= 0) then
        write(*,*) 'psdm_rtm: no usable HIP device (psdm_rtm_create returned ', rc, ')'
        error stop
      endif
    endif
    psdm_context = psdm_ctx
  end function

  subroutine psdm_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(psdm_rtm_last_error(psdm_ctx), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    enddo
    write(*,*) 'psdm_rtm: ', where, ' failed (', rc, '): ', msg(1:n)
    error stop
  end subroutine

end module psdm_rtm_iface


subroutine solver_ac_forward_ram(rank,NumSou,NumRec,nxSou,nySou,nxRec,nyRec,&
S,dmodel,dt,ny,nx,ctmp,&
store_snap,nt_snap,nt,Fields,Rec_point)
  use mod_parfile, only: dPML, fs, taper_option, T_taper
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt
  integer :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
  real :: dmodel, dt
  real :: S(nt,NumSou), ctmp(ny,nx), Fields(ny,nx,nt_snap), Rec_point(nt,NumRec)
  integer(c_int) :: N_taper
  call psdm_check(psdm_rtm_forward_ram(psdm_context(), dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, &
                  dmodel, dt, ny, nx, ctmp, store_snap, nt_snap, nt, Fields, 0, Rec_point), 'solver_ac_forward_ram')
  if (taper_option /= 0) then
    N_taper = max(1, int(T_taper / dt))
    call psdm_check(psdm_rtm_taper_snapshots(psdm_context(), Fields, nt, N_taper), 'solver_ac_forward_ram (taper)')
  endif
end subroutine


subroutine solver_ac_backward_ram(rank,NumSou,NumRec,nxSou,nySou,nxRec,nyRec,&
Sadj,dmodel,dt,ny,nx,ctmp,&
store_snap,nt_snap,nt,Fields,kernel)
  use mod_parfile, only: dPML, fs
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt
  integer :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
  real :: dmodel, dt
  real :: Sadj(nt,NumRec), ctmp(ny,nx), Fields(ny,nx,nt_snap), kernel(ny,nx)
  call psdm_check(psdm_rtm_backward_ram(psdm_context(), dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, &
                  dmodel, dt, ny, nx, ctmp, store_snap, nt_snap, nt, Fields, 0, kernel), 'solver_ac_backward_ram')
end subroutine


subroutine solver_ac_forward_disk(rank,NumSou,NumRec,nxSou,nySou,nxRec,nyRec,&
S,dmodel,dt,ny,nx,ctmp,&
shotid,unit_fields,folder_fields,&
store_snap,nt_snap,nt,Rec_point)
  use mod_parfile, only: dPML, fs
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, shotid, unit_fields
  integer :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
  real :: dmodel, dt
  real :: S(nt,NumSou), ctmp(ny,nx), Rec_point(nt,NumRec)
  character(len=500) :: folder_fields
  ! the reference writes one file per snapshot under folder_fields; here they stay in HBM
  call psdm_check(psdm_rtm_forward_disk(psdm_context(), dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, &
                  dmodel, dt, ny, nx, ctmp, shotid, store_snap, nt_snap, nt, Rec_point), 'solver_ac_forward_disk')
end subroutine


subroutine solver_ac_backward_disk(rank,NumSou,NumRec,nxSou,nySou,nxRec,nyRec,&
Sadj,dmodel,dt,ny,nx,ctmp,&
shotid,unit_fields,folder_fields,&
store_snap,nt_snap,nt,kernel)
  use mod_parfile, only: dPML, fs
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, shotid, unit_fields
  integer :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
  real :: dmodel, dt
  real :: Sadj(nt,NumRec), ctmp(ny,nx), kernel(ny,nx)
  character(len=500) :: folder_fields
  call psdm_check(psdm_rtm_backward_disk(psdm_context(), dPML, fs, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, &
                  dmodel, dt, ny, nx, ctmp, shotid, store_snap, nt_snap, nt, kernel), 'solver_ac_backward_disk')
end subroutine


! Gather conditioning of the driver on the device (calculate_PSDM_RTM.f90:299-358): the
! frequency filter (when f_init > 0 and f_final > 0), the resampling to the propagation
! step (or the copy of the first nt_prop2 samples), the cut_sf zeroing and the taper /
! cut_multiple of the trace ends, with the driver's own variables.  A caller replaces that
! block with one call.  Field_Data is left as it was (the reference filters it in place;
! the driver does not read it again for this shot).
subroutine psdm_condition_gather(Field_Data, nt_data, dt_data, NumRec, typef, f_init, f_final, &
    Data_target, nt_prop2, dt_prop, nt_taper, taper_option, N_taper, N_zeros, cut_sf, nt_sf, cut_multiple, nt_cut)
  use psdm_rtm_iface
  implicit none
  integer :: nt_data, NumRec, typef, nt_prop2, nt_taper, taper_option, N_taper, N_zeros, cut_sf, cut_multiple
  real :: dt_data, f_init, f_final, dt_prop
  real :: Field_Data(nt_data, NumRec), Data_target(nt_taper, NumRec)
  integer(c_int), target :: nt_sf(NumRec), nt_cut(NumRec)
  type(psdm_rtm_cond) :: q
  q%nt_data = nt_data
  q%dt_data = dt_data
  q%typef = typef
  q%f_init = f_init
  q%f_final = f_final
  q%nt_prop2 = nt_prop2
  q%dt_prop = dt_prop
  q%nt_taper = nt_taper
  q%taper_option = taper_option
  q%N_taper = N_taper
  q%N_zeros = N_zeros
  q%nt_sf = c_null_ptr
  q%nt_cut = c_null_ptr
  if (cut_sf > 0) q%nt_sf = c_loc(nt_sf)
  if (cut_multiple > 0) q%nt_cut = c_loc(nt_cut)
  call psdm_check(psdm_rtm_condition_host(psdm_context(), q, NumRec, Field_Data, Data_target), &
                  'psdm_condition_gather')
end subroutine

! Replaces get_shot_data (get_mcs_data.f90:150-214) with its own arguments, after
! open_shot_mcs (:286-314) would have opened the shot's file: the file is
! folder_input_mcs // su_file_mcs(icount), NumRec_MCS traces of nt_mcs_max samples, byte
! order endianness_data (1 big-endian, 0 little).  raw_data gets the first nt_mcs samples
! of each trace; with delay_time_data /= 0, delay_data gets them shifted by
! int(delay_time_data/dt_mcs) samples (0 where nothing is shifted in).
subroutine psdm_get_shot_data(icount, raw_data, delay_data)
  use psdm_rtm_iface
  use mod_parfile, only: folder_input_mcs, su_file_mcs, endianness_data, NumRec_MCS, nt_mcs, nt_mcs_max, &
                         delay_time_data, dt_mcs
  implicit none
  integer :: icount
  real :: raw_data(nt_mcs, NumRec_MCS), delay_data(nt_mcs, NumRec_MCS)
  character(len=1000) :: file_name
  integer :: id
  file_name = trim(adjustl(folder_input_mcs)) // trim(adjustl(su_file_mcs(icount)))
  call psdm_check(psdm_rtm_su_read_host(psdm_context(), trim(file_name) // c_null_char, 0_c_long_long, NumRec_MCS, &
                                        nt_mcs_max, nt_mcs, endianness_data, 0_c_int, raw_data), 'psdm_get_shot_data')
  if (delay_time_data /= 0.) then
    id = int(delay_time_data / dt_mcs)
    call psdm_check(psdm_rtm_su_read_host(psdm_context(), trim(file_name) // c_null_char, 0_c_long_long, &
                                          NumRec_MCS, nt_mcs_max, nt_mcs, endianness_data, id, delay_data), &
                    'psdm_get_shot_data')
  endif
end subroutine

! Replace dynamic_window_tt_forward / dynamic_window_tt_backward (dynamic_window.f90:230-325)
! with their own argument lists: the window of the shot from the reference's dynamic_window,
! then the travel times of the source (TT_s) or of all receivers at once (TT_r) on the
! device, bit for bit those of solve_eikonal_FSM.
subroutine psdm_window_tt_forward(rank,NumRec,nxSou,nySou,nxRec,nyRec,dd,ny,nx,nx_w,Vp_w,dd_space,TT_s)
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumRec, ny, nx, nx_w, x_i, x_f
  integer :: nxSou(1), nySou(1), nxRec(NumRec), nyRec(NumRec)
  real :: dd, dd_space
  real :: Vp_w(ny, nx_w), TT_s(ny, nx_w)
  integer(c_int) :: nxs(1), nys(1)
  call dynamic_window(rank, 1, NumRec, nxSou, nxRec, dd, nx, dd_space, nx_w, x_i, x_f)
  nxs = nxSou - x_i + 1
  nys = nySou
  call psdm_check(psdm_rtm_eikonal_host(psdm_context(), Vp_w, ny, nx_w, dd, dd, 1_c_int, nxs, nys, TT_s), &
                  'psdm_window_tt_forward')
end subroutine

subroutine psdm_window_tt_backward(rank,NumRec,nxSou,nySou,nxRec,nyRec,dd,ny,nx,nx_w,Vp_w,dd_space,TT_r)
  use psdm_rtm_iface
  implicit none
  integer :: rank, NumRec, ny, nx, nx_w, x_i, x_f, j
  integer :: nxSou(1), nySou(1), nxRec(NumRec), nyRec(NumRec)
  real :: dd, dd_space
  real :: Vp_w(ny, nx_w), TT_r(ny, nx_w, NumRec)
  integer(c_int), allocatable :: nxr(:), nyr(:)
  call dynamic_window(rank, 1, NumRec, nxSou, nxRec, dd, nx, dd_space, nx_w, x_i, x_f)
  allocate(nxr(NumRec), nyr(NumRec))
  nxr = nxRec - x_i + 1
  nyr = nyRec
  do j = 1, NumRec  ! the reference stops on receivers outside the window (:307-315)
    if (nxr(j) < 1 .or. nxr(j) > nx_w) then
      write(*,*) "nxr out of bounds"
      write(*,*) j, nxr(j), nx_w
      stop
    endif
    if (nyr(j) < 1 .or. nyr(j) > ny) stop "nyr out of bounds"
  enddo
  call psdm_check(psdm_rtm_eikonal_host(psdm_context(), Vp_w, ny, nx_w, dd, dd, NumRec, nxr, nyr, TT_r), &
                  'psdm_window_tt_backward')
  deallocate(nxr, nyr)
end subroutine
