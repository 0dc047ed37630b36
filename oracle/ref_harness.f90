! C-callable entry points onto the reference Fortran solvers -- TEST INFRASTRUCTURE ONLY.
!
! Linked by oracle/Makefile together with the reference's own, unmodified sources
! (src/modules.f90 and the six solver files, compiled where they lie), into
! oracle/_ref/libpsdm_ref.so.  Used by tests/golden/make_golden.py to write golden
! vectors and by bench.py's cpu_baseline leg ("kind": "reference").
!
! Each wrapper sets the two module switches the solvers read (mod_parfile: dPML, fs;
! solver_ac_forward_ram.f90:8) and calls the solver with the reference's own
! argument list (solver_ac_forward_ram.f90:4-6, solver_ac_backward_ram.f90:4-6,
! solver_ac_forward_disk.f90:4-7, solver_ac_backward_disk.f90:4-7).
module ref_harness
  use iso_c_binding
  use mod_parfile, only: dPML, fs
  implicit none
contains

  subroutine ref_forward_ram(dpml_in, fs_in, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, dmodel, dt, &
                             ny, nx, ctmp, store_snap, nt_snap, nt, Fields, Rec_point) bind(C, name="ref_forward_ram")
    integer(c_int), value :: dpml_in, fs_in, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt
    integer(c_int) :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
    real(c_float), value :: dmodel, dt
    real(c_float) :: S(nt, NumSou), ctmp(ny, nx), Fields(ny, nx, nt_snap), Rec_point(nt, NumRec)
    integer :: rank, ns, nr, ny_, nx_, ss, nts, nt_
    real :: dm, dtt
    rank = 0; ns = NumSou; nr = NumRec; ny_ = ny; nx_ = nx; ss = store_snap; nts = nt_snap; nt_ = nt
    dm = dmodel; dtt = dt
    dPML = dpml_in; fs = fs_in
    call solver_ac_forward_ram(rank, ns, nr, nxSou, nySou, nxRec, nyRec, S, dm, dtt, ny_, nx_, ctmp, &
                               ss, nts, nt_, Fields, Rec_point)
  end subroutine

  subroutine ref_backward_ram(dpml_in, fs_in, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, &
                              ny, nx, ctmp, store_snap, nt_snap, nt, Fields, kernel) bind(C, name="ref_backward_ram")
    integer(c_int), value :: dpml_in, fs_in, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt
    integer(c_int) :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
    real(c_float), value :: dmodel, dt
    real(c_float) :: Sadj(nt, NumRec), ctmp(ny, nx), Fields(ny, nx, nt_snap), kernel(ny, nx)
    integer :: rank, ns, nr, ny_, nx_, ss, nts, nt_
    real :: dm, dtt
    rank = 0; ns = NumSou; nr = NumRec; ny_ = ny; nx_ = nx; ss = store_snap; nts = nt_snap; nt_ = nt
    dm = dmodel; dtt = dt
    dPML = dpml_in; fs = fs_in
    call solver_ac_backward_ram(rank, ns, nr, nxSou, nySou, nxRec, nyRec, Sadj, dm, dtt, ny_, nx_, ctmp, &
                                ss, nts, nt_, Fields, kernel)
  end subroutine

  subroutine c_to_folder(cstr, folder)
    character(kind=c_char) :: cstr(*)
    character(len=500) :: folder
    integer :: n
    folder = ' '
    n = 0
    do while (cstr(n + 1) /= c_null_char .and. n < 499)
      n = n + 1
      folder(n:n) = cstr(n)
    enddo
  end subroutine

  subroutine ref_forward_disk(dpml_in, fs_in, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, dmodel, dt, &
                              ny, nx, ctmp, shotid, folder_c, store_snap, nt_snap, nt, Rec_point) &
                              bind(C, name="ref_forward_disk")
    integer(c_int), value :: dpml_in, fs_in, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, shotid
    integer(c_int) :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
    real(c_float), value :: dmodel, dt
    real(c_float) :: S(nt, NumSou), ctmp(ny, nx), Rec_point(nt, NumRec)
    character(kind=c_char) :: folder_c(*)
    character(len=500) :: folder
    integer :: rank, ns, nr, ny_, nx_, ss, nts, nt_, sid, unit_f
    real :: dm, dtt
    call c_to_folder(folder_c, folder)
    rank = 0; ns = NumSou; nr = NumRec; ny_ = ny; nx_ = nx; ss = store_snap; nts = nt_snap; nt_ = nt
    sid = shotid; unit_f = 77; dm = dmodel; dtt = dt
    dPML = dpml_in; fs = fs_in
    call solver_ac_forward_disk(rank, ns, nr, nxSou, nySou, nxRec, nyRec, S, dm, dtt, ny_, nx_, ctmp, &
                                sid, unit_f, folder, ss, nts, nt_, Rec_point)
  end subroutine

  subroutine ref_backward_disk(dpml_in, fs_in, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, &
                               ny, nx, ctmp, shotid, folder_c, store_snap, nt_snap, nt, kernel) &
                               bind(C, name="ref_backward_disk")
    integer(c_int), value :: dpml_in, fs_in, NumSou, NumRec, ny, nx, store_snap, nt_snap, nt, shotid
    integer(c_int) :: nxSou(NumSou), nySou(NumSou), nxRec(NumRec), nyRec(NumRec)
    real(c_float), value :: dmodel, dt
    real(c_float) :: Sadj(nt, NumRec), ctmp(ny, nx), kernel(ny, nx)
    character(kind=c_char) :: folder_c(*)
    character(len=500) :: folder
    integer :: rank, ns, nr, ny_, nx_, ss, nts, nt_, sid, unit_f
    real :: dm, dtt
    call c_to_folder(folder_c, folder)
    rank = 0; ns = NumSou; nr = NumRec; ny_ = ny; nx_ = nx; ss = store_snap; nts = nt_snap; nt_ = nt
    sid = shotid; unit_f = 77; dm = dmodel; dtt = dt
    dPML = dpml_in; fs = fs_in
    call solver_ac_backward_disk(rank, ns, nr, nxSou, nySou, nxRec, nyRec, Sadj, dm, dtt, ny_, nx_, ctmp, &
                                 sid, unit_f, folder, ss, nts, nt_, kernel)
  end subroutine

  ! PML_coeffs (solver_PML_coeffs.f90:1-75): fA,fB(dn), corners ld,rd,lu,ru (dn,dn) each
  subroutine ref_pml_coeffs(dpml_in, dt, d, fA, fB, cA, cB) bind(C, name="ref_pml_coeffs")
    integer(c_int), value :: dpml_in
    real(c_float), value :: dt, d
    real(c_float) :: fA(dpml_in + 6), fB(dpml_in + 6), cA(dpml_in + 6, dpml_in + 6, 4), cB(dpml_in + 6, dpml_in + 6, 4)
    integer :: acc, dp, dn
    real :: dtt, dd
    acc = 3; dp = dpml_in; dn = dp + 2 * acc; dtt = dt; dd = d
    fA = 0; fB = 0; cA = 0; cB = 0
    call PML_coeffs(acc, dp, dtt, dd, dn, fA, fB, cA(:, :, 1), cA(:, :, 2), cA(:, :, 3), cA(:, :, 4), &
                    cB(:, :, 1), cB(:, :, 2), cB(:, :, 3), cB(:, :, 4))
  end subroutine

  ! finite_difference_coef (solver_FD_coeffs.f90:1-73), accuracy 3
  subroutine ref_fd_coef(num, coef) bind(C, name="ref_fd_coef")
    integer(c_int), value :: num
    real(c_float) :: coef(4)
    integer :: n, acc
    n = num; acc = 3
    coef = 0
    call finite_difference_coef(n, acc, coef)
  end subroutine

  ! space_deriv2x / 2y / derivx / derivy (solver_space_derivs_clean.f90:1-140), acc = 3
  subroutine ref_deriv(which, nyt, nxt, p, d) bind(C, name="ref_deriv")
    integer(c_int), value :: which, nyt, nxt
    real(c_float) :: p(nyt, nxt), d(nyt, nxt)
    real :: c1(4), c2(4)
    integer :: acc, ny_, nx_
    acc = 3; ny_ = nyt; nx_ = nxt
    call finite_difference_coef(1, acc, c1)
    call finite_difference_coef(2, acc, c2)
    select case (which)
    case (0)
      call space_deriv2x(acc, c2, ny_, nx_, p, d)
    case (1)
      call space_deriv2y(acc, c2, ny_, nx_, p, d)
    case (2)
      call space_derivx(acc, c1, ny_, nx_, p, d)
    case default
      call space_derivy(acc, c1, ny_, nx_, p, d)
    end select
  end subroutine

  ! Data conditioning of a gather before migration (calculate_PSDM_RTM.f90:299-358): the
  ! reference's own frequency filter (auxiliar.f90:498-531 -> filters.f90:50-234 with
  ! FFTPACK), Akima resampling (interpolation_akima.f90:1-31 -> akima433.f90) and the
  ! half-cosine taper with zero padding (filters.f90:20-48).
  subroutine ref_freq_filter(nt, dt, Num, typef, f1, f2, data) bind(C, name="ref_freq_filter")
    integer(c_int), value :: nt, Num, typef
    real(c_float), value :: dt, f1, f2
    real(c_float) :: data(nt, Num)
    integer :: nt_, num_, tf
    real :: dt_, fa, fb
    nt_ = nt; num_ = Num; tf = typef; dt_ = dt; fa = f1; fb = f2
    call freq_filter(data, nt_, dt_, num_, tf, fa, fb)
  end subroutine

  subroutine ref_interpolation_akima(nt_data, dt_data, NumRec, data, nt, dt, out) &
      bind(C, name="ref_interpolation_akima")
    integer(c_int), value :: nt_data, NumRec, nt
    real(c_float), value :: dt_data, dt
    real(c_float) :: data(nt_data, NumRec), out(nt, NumRec)
    logical, external :: interpolation_akima
    logical :: ok
    integer :: ntd, nr, ntt
    real :: dtd, dtt
    ntd = nt_data; nr = NumRec; ntt = nt; dtd = dt_data; dtt = dt
    ok = interpolation_akima(data, dtd, ntd, out, nr, dtt, ntt)
  end subroutine

  subroutine ref_taper_pad_trace(trace, nt_prop, N_taper, N_zeros, nt_new) bind(C, name="ref_taper_pad_trace")
    integer(c_int), value :: nt_prop, N_taper, N_zeros, nt_new
    real(c_float) :: trace(nt_new)
    integer :: a, b, c, d
    a = nt_prop; b = N_taper; c = N_zeros; d = nt_new
    call taper_pad_trace(trace, a, b, c, d)
  end subroutine

  ! The linear interpolation of the depth-to-time conversion (psdm2twt,
  ! auxiliar_PSDM.f90:23-91 -> linear_interpolation_1D, linear_interpolation.f90:72-103).
  subroutine ref_linear_interpolation_1d(x1, y1, n1, x2, y2, n2) bind(C, name="ref_linear_interpolation_1d")
    integer(c_int), value :: n1, n2
    real(c_float) :: x1(n1), y1(n1), x2(n2), y2(n2)
    integer :: a, b
    a = n1; b = n2
    call linear_interpolation_1D(x1, y1, a, x2, y2, b)
  end subroutine

  ! The SU trace reader (open_shot_mcs get_mcs_data.f90:286-314, get_shot_data :150-214):
  ! nrec traces of ntmax samples behind 60-word headers, the first nt samples of each.
  subroutine ref_su_read_shot(cpath, big, nrec, ntmax, nt, raw) bind(C, name="ref_su_read_shot")
    use mod_parfile, only: delay_time_data, dt_mcs
    character(kind=c_char) :: cpath(*)
    integer(c_int), value :: big, nrec, ntmax, nt
    real(c_float) :: raw(nt, nrec)
    real, allocatable :: delay(:, :)
    call su_setup(cpath, big, nrec, ntmax, nt)
    ! get_shot_data tests an uninitialised local before computing the delay shift; with
    ! delay_time_data = 0 that shift is 0 whatever the test decides
    delay_time_data = 0.; dt_mcs = 1.
    allocate(delay(nt, nrec))
    delay = 0.
    call open_shot_mcs(0, 1)
    call get_shot_data(1, raw, delay)
    call close_mcs_files(1)
    deallocate(delay)
  end subroutine

  subroutine su_setup(cpath, big, nrec, ntmax, nt)
    use mod_parfile, only: folder_input_mcs, su_file_mcs, endianness_data, endianness_machine, NumRec_MCS, &
                           nt_mcs, nt_mcs_max
    character(kind=c_char) :: cpath(*)
    integer(c_int), value :: big, nrec, ntmax, nt
    character(len=500) :: p
    call c_to_folder(cpath, p)
    folder_input_mcs = ' '
    if (allocated(su_file_mcs)) deallocate(su_file_mcs)
    allocate(su_file_mcs(1))
    su_file_mcs(1) = trim(p)
    endianness_machine = 0; endianness_data = big
    NumRec_MCS = nrec; nt_mcs = nt; nt_mcs_max = ntmax
  end subroutine

  ! Header words and receiver geometry (get_pos_channels get_mcs_data.f90:418-510,
  ! get_geometry_mcs :128-148) of one shot file.
  subroutine ref_su_geometry(cpath, big, nrec, ntmax, byte_shot, offset_hdr, offset_unit_in, dir, near_off, &
                             drec_in, pos_shot, dd, shot_depth, rec_depth, shot_id, offs, posrec, nxs, nys, &
                             nxr, nyr) bind(C, name="ref_su_geometry")
    use mod_parfile, only: byte_shotnumber_mcs, offset_header, offset_unit, dir_boat, near_offset, drec, &
                           shot_depth_mcs, streamer_depth
    use mod_data_arrays, only: shotID_su_mcs, offset_su, pos_rec, pos_shot_mcs, nxSou_MCS, nySou_MCS, &
                               nxRec_MCS, nyRec_MCS
    include 'mpif.h'
    character(kind=c_char) :: cpath(*)
    integer(c_int), value :: big, nrec, ntmax, byte_shot, offset_hdr, dir
    real(c_float), value :: offset_unit_in, near_off, drec_in, pos_shot, dd, shot_depth, rec_depth
    integer(c_int) :: shot_id, nxs, nys, nxr(nrec), nyr(nrec)
    real(c_float) :: offs(nrec), posrec(nrec)
    integer :: ierr
    logical :: up
    real :: d
    call MPI_Initialized(up, ierr)
    if (.not. up) call MPI_Init(ierr)
    call su_setup(cpath, big, nrec, ntmax, ntmax)
    byte_shotnumber_mcs = byte_shot; offset_header = offset_hdr; offset_unit = offset_unit_in
    dir_boat = dir; near_offset = near_off; drec = drec_in
    shot_depth_mcs = shot_depth; streamer_depth = rec_depth
    if (allocated(shotID_su_mcs)) deallocate(shotID_su_mcs, offset_su, pos_rec, pos_shot_mcs, nxSou_MCS, &
                                             nySou_MCS, nxRec_MCS, nyRec_MCS)
    allocate(shotID_su_mcs(1), offset_su(nrec, 1), pos_rec(nrec, 1), pos_shot_mcs(1), nxSou_MCS(1), &
             nySou_MCS(1), nxRec_MCS(nrec, 1), nyRec_MCS(nrec, 1))
    offset_su = 0.; pos_shot_mcs(1) = pos_shot
    call open_shot_mcs(0, 1)
    call get_pos_channels(0, 1)
    d = dd
    call get_geometry_mcs(0, 1, d)
    call close_mcs_files(1)
    shot_id = shotID_su_mcs(1); nxs = nxSou_MCS(1); nys = nySou_MCS(1)
    offs = offset_su(:, 1); posrec = pos_rec(:, 1); nxr = nxRec_MCS(:, 1); nyr = nyRec_MCS(:, 1)
  end subroutine

  ! Shot positions along the line from the navigation file (read_nav_shot_mcs
  ! get_mcs_data.f90:30-126), called as a rank other than 0 (no report, no stop).
  subroutine ref_read_nav(cpath, cols, dir, add, n, ids, pos, bat, li, lf) bind(C, name="ref_read_nav")
    use mod_parfile, only: folder_input_mcs, nav_shot_mcs, nav_mcs_cols, dir_boat, shot_i, shot_f, &
                           d_shot_model, line_i, line_f
    use mod_data_arrays, only: shotID_nav_mcs, pos_shot_mcs, pos_bat_mcs
    character(kind=c_char) :: cpath(*)
    integer(c_int), value :: cols, dir, n
    real(c_float), value :: add
    integer(c_int) :: ids(n), li, lf
    real(c_float) :: pos(n), bat(n)
    character(len=500) :: p
    call c_to_folder(cpath, p)
    folder_input_mcs = ' '; nav_shot_mcs = trim(p); nav_mcs_cols = cols; dir_boat = dir
    d_shot_model = add; shot_i = 0; shot_f = 0; line_i = 0; line_f = 0
    if (allocated(shotID_nav_mcs)) deallocate(shotID_nav_mcs)
    if (allocated(pos_shot_mcs)) deallocate(pos_shot_mcs)
    if (allocated(pos_bat_mcs)) deallocate(pos_bat_mcs)
    allocate(shotID_nav_mcs(n), pos_shot_mcs(n), pos_bat_mcs(n))
    call read_nav_shot_mcs(1)
    ids = shotID_nav_mcs; pos = pos_shot_mcs; bat = pos_bat_mcs; li = line_i; lf = line_f
  end subroutine

  ! get_su_data (auxiliar.f90:9-44) on a file opened as get_vel opens an SU model
  ! (:92, CONVERT='BIG_ENDIAN'), or little-endian: ni traces of ns samples, the first
  ! ns_max of each.
  subroutine ref_get_su_data(cpath, big, ni, ns, ns_max, out) bind(C, name="ref_get_su_data")
    character(kind=c_char) :: cpath(*)
    integer(c_int), value :: big, ni, ns, ns_max
    real(c_float) :: out(ns_max, ni)
    character(len=500) :: p
    integer :: a, b, c, u
    call c_to_folder(cpath, p)
    u = 77
    if (big /= 0) then
      open(u, FILE=trim(p), ACCESS='stream', FORM='unformatted', CONVERT='BIG_ENDIAN', STATUS='old')
    else
      open(u, FILE=trim(p), ACCESS='stream', FORM='unformatted', CONVERT='LITTLE_ENDIAN', STATUS='old')
    endif
    a = ni; b = ns; c = ns_max
    call get_su_data(a, b, c, u, out)
    close(u)
  end subroutine

  ! solve_eikonal_FSM (solve_eikonal_FSM.f90:1-83) for one source point (1-based sx, sz).
  subroutine ref_eikonal(nx, ny, dx, dy, v, sx, sz, T) bind(C, name="ref_eikonal")
    integer(c_int), value :: nx, ny, sx, sz
    real(c_float), value :: dx, dy
    real(c_float) :: v(ny, nx), T(ny, nx)
    integer :: a, b, px(1), pz(1)
    real :: ddx, ddy
    a = nx; b = ny; px(1) = sx; pz(1) = sz; ddx = dx; ddy = dy
    call solve_eikonal_FSM(a, b, ddx, ddy, v, px, pz, T)
  end subroutine

  ! normalize_psdm (auxiliar_PSDM.f90:1-20): PSDM = PSDM_sum / Cov_sum where Cov_sum > 0,
  ! else 0, in the reference's FP32.
  subroutine ref_normalize_psdm(nx, ny, psdm_sum, cov_sum, psdm) bind(C, name="ref_normalize_psdm")
    integer(c_int), value :: nx, ny
    real(c_float) :: psdm_sum(ny, nx), cov_sum(ny, nx), psdm(ny, nx)
    integer :: a, b
    real :: dx, dy
    a = nx; b = ny; dx = 1.0; dy = 1.0
    call normalize_psdm(a, b, dx, dy, psdm_sum, cov_sum, psdm)
  end subroutine

  ! Ricker_mcs (get_mcs_data.f90:318-351): the wavelet reads nt_prop, dt_prop,
  ! freq_ricker and delay_ricker from mod_parfile.
  subroutine ref_ricker_mcs(nt, dt, f0, t0, s0) bind(C, name="ref_ricker_mcs")
    use mod_parfile, only: nt_prop, dt_prop, freq_ricker, delay_ricker
    integer(c_int), value :: nt
    real(c_float), value :: dt, f0, t0
    real(c_float) :: s0(nt)
    integer :: rank
    nt_prop = nt; dt_prop = dt; freq_ricker = f0; delay_ricker = t0; rank = 0
    call Ricker_mcs(rank, s0)
  end subroutine

end module ref_harness
