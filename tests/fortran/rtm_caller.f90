! Test caller of the Fortran drop-in: reproduces the solver calls of one RTM shot as
! calculate_PSDM_RTM.f90:395-466 makes them (Fields allocated and zeroed by the caller,
! forward, BackSource = data reversed in time, PSDM_w zeroed, backward with
! nt_snap = nt), going through psdm_rtm_fortran.f90 -> libpsdm_rtm.so.  With taper != 0
! the driver would next taper its host Fields (calculate_PSDM_RTM.f90:421-427); the shim
! keeps the snapshots on the device and tapers them there, so the host Fields stay zero
! and their taper changes nothing: it is left out here.
! Usage: rtm_caller in.bin out.bin
!   in.bin : int32 ny,nx,nt,NumRec,dPML,fs,taper,variant; float32 dmodel,dt,T_taper;
!            int32 nxSou,nySou,nxRec(NumRec),nyRec(NumRec); float32 S(nt), ctmp(ny,nx),
!            Data(nt,NumRec)
!   out.bin: float32 Rec_point(nt,NumRec), kernel(ny,nx)
program rtm_caller
  use mod_parfile, only: dPML, fs, taper_option, T_taper
  implicit none
  integer :: ny, nx, nt, NumRec, variant, u, ierr
  integer :: nxSou(1), nySou(1)
  integer, allocatable :: nxRec(:), nyRec(:)
  real :: dmodel, dt
  real, allocatable :: S(:,:), ctmp(:,:), Data(:,:), Rec(:,:), BackSource(:,:), Fields(:,:,:), kernel(:,:)
  character(len=500) :: fin, fout, folder
  integer :: j

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ierr)
  if (ierr /= 0) error stop 'cannot open input'
  read(u) ny, nx, nt, NumRec, dPML, fs, taper_option, variant
  read(u) dmodel, dt, T_taper
  allocate(nxRec(NumRec), nyRec(NumRec))
  read(u) nxSou, nySou, nxRec, nyRec
  allocate(S(nt,1), ctmp(ny,nx), Data(nt,NumRec), Rec(nt,NumRec), BackSource(nt,NumRec), kernel(ny,nx))
  read(u) S, ctmp, Data
  close(u)

  Rec = 0
  folder = 'FIELDS/'
  if (variant == 0) then
    allocate(Fields(ny,nx,nt))
    Fields = 0
    call solver_ac_forward_ram(0,1,NumRec,nxSou,nySou,nxRec,nyRec,S,dmodel,dt,ny,nx,ctmp,1,nt,nt,Fields,Rec)
  else
    call solver_ac_forward_disk(0,1,NumRec,nxSou,nySou,nxRec,nyRec,S,dmodel,dt,ny,nx,ctmp,11,500,folder,1,nt,nt,Rec)
  endif
  do j = 1, NumRec
    BackSource(:,j) = Data(nt:1:-1,j)
  enddo
  kernel = 0
  if (variant == 0) then
    call solver_ac_backward_ram(0,1,NumRec,nxSou,nySou,nxRec,nyRec,BackSource,dmodel,dt,ny,nx,ctmp,1,nt,nt,Fields,kernel)
  else
    call solver_ac_backward_disk(0,1,NumRec,nxSou,nySou,nxRec,nyRec,BackSource,dmodel,dt,ny,nx,ctmp,11,500,folder, &
                                 1,nt,nt,kernel)
  endif

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) Rec, kernel
  close(u)
end program rtm_caller
