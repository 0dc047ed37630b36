/*
 * psdm_rtm.h -- C ABI of the MI355X (gfx950) RTM solver library libpsdm_rtm.so.
 *
 * Drop-in for the four time-stepping solvers of PSDM's RTM path.  The reference calls
 * them from dynamic_window.f90:90-92,134-137,177-179,221-224; the Fortran iso_c_binding
 * shim psdm_amd/csrc/psdm_rtm_fortran.f90 defines subroutines with the reference's
 * names and argument lists and forwards to the entry points below, so
 * calculate_PSDM_RTM.f90 and dynamic_window.f90 link unchanged.
 *
 * Conventions (as in the reference): arrays are column-major with depth fastest;
 * geometry indices are 1-based grid indices inside the shot window; scalars are
 * int32 / float32.  Every entry point returns 0 on success and a negative code on
 * error; psdm_rtm_last_error() then describes the failure.  A context is not
 * re-entrant: use one per host thread.
 */
#ifndef PSDM_RTM_H
#define PSDM_RTM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psdm_rtm_ctx psdm_rtm_ctx;

enum {
    PSDM_RTM_OK = 0,
    PSDM_RTM_EINVAL = -1,  /* argument outside the solver's defined range */
    PSDM_RTM_EDEVICE = -2, /* HIP runtime error (allocation, launch) */
    PSDM_RTM_ENOENT = -3,  /* no device-resident snapshots under that key */
};

/* Context on HIP device `device`; owns the device buffers and one stream. */
int psdm_rtm_create(psdm_rtm_ctx **ctx, int device);
int psdm_rtm_destroy(psdm_rtm_ctx *ctx);
const char *psdm_rtm_last_error(const psdm_rtm_ctx *ctx);

/*
 * Replaces solver_ac_forward_ram (solver_ac_forward_ram.f90:4-6).
 * dPML, fs: the mod_parfile values the reference reads (solver_ac_forward_ram.f90:8).
 * S(nt,NumSou); ctmp(ny,nx); Rec_point(nt,NumRec) is overwritten.
 * Fields(ny,nx,nt_snap): host_fields != 0 -> snapshots are copied into Fields (needed
 * when the caller tapers them, calculate_PSDM_RTM.f90:421-427); host_fields == 0 ->
 * snapshots stay in HBM, keyed by the Fields pointer, for the matching backward call.
 */
int psdm_rtm_forward_ram(psdm_rtm_ctx *ctx, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                         const int *nySou, const int *nxRec, const int *nyRec, const float *S, float dmodel,
                         float dt, int ny, int nx, const float *ctmp, int store_snap, int nt_snap, int nt,
                         float *Fields, int host_fields, float *Rec_point);

/*
 * Replaces solver_ac_backward_ram (solver_ac_backward_ram.f90:4-6).
 * Sadj(nt,NumRec); kernel(ny,nx) is accumulated (+=) as in the reference.
 * host_fields != 0: snapshots are read from Fields; == 0: from HBM under the key
 * Fields left by psdm_rtm_forward_ram, which are released afterwards.
 */
int psdm_rtm_backward_ram(psdm_rtm_ctx *ctx, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                          const int *nySou, const int *nxRec, const int *nyRec, const float *Sadj, float dmodel,
                          float dt, int ny, int nx, const float *ctmp, int store_snap, int nt_snap, int nt,
                          const float *Fields, int host_fields, float *kernel);

/*
 * Replace solver_ac_forward_disk / solver_ac_backward_disk (solver_ac_forward_disk.f90:4-7,
 * solver_ac_backward_disk.f90:4-7).  The per-step snapshot files of the reference
 * become device-resident snapshots keyed by shotid; the backward call consumes and
 * frees them.  The _disk stepping order (strips scattered after the free-surface row
 * is zeroed, full-width velocity pads) is kept.
 */
int psdm_rtm_forward_disk(psdm_rtm_ctx *ctx, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                          const int *nySou, const int *nxRec, const int *nyRec, const float *S, float dmodel,
                          float dt, int ny, int nx, const float *ctmp, int shotid, int store_snap, int nt_snap,
                          int nt, float *Rec_point);
int psdm_rtm_backward_disk(psdm_rtm_ctx *ctx, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                           const int *nySou, const int *nxRec, const int *nyRec, const float *Sadj, float dmodel,
                           float dt, int ny, int nx, const float *ctmp, int shotid, int store_snap, int nt_snap,
                           int nt, float *kernel);

/* ------------------------------------------------------------------------------
 * Device-resident form (no counterpart in the reference: the one-shot drivers above
 * and the batched engine are built on it).  All float pointers are device pointers
 * on the context's device; geometry arrays are host pointers.
 * ------------------------------------------------------------------------------ */
typedef struct {
    int ny, nx, nt;          /* shot window (depth samples, columns), time steps */
    int dPML, fs;            /* CPML width, free surface flag (mod_parfile) */
    int variant;             /* 0 = _ram stepping semantics, 1 = _disk */
    int store_snap, nt_snap; /* snapshot interval and count, as in the reference */
    float dmodel, dt;
    int NumSou, NumRec;
    const int *nxSou, *nySou; /* forward injection points (1-based) */
    const int *nxRec, *nyRec; /* receivers (1-based) */
} psdm_rtm_geom;

/* Forward on device: d_ctmp(ny,nx), d_S(nt,NumSou) -> d_rec(nt,NumRec); snapshots kept
 * in HBM under `key`.  Launches are ordered on the context stream; the call returns
 * once they are queued (psdm_rtm_sync waits). */
int psdm_rtm_forward_device(psdm_rtm_ctx *ctx, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_S,
                            float *d_rec, long long key);
/* Backward on device: d_Sadj(nt,NumRec), snapshots under `key` (released afterwards),
 * d_kernel(ny,nx) accumulated. */
int psdm_rtm_backward_device(psdm_rtm_ctx *ctx, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_Sadj,
                             float *d_kernel, long long key);
int psdm_rtm_sync(psdm_rtm_ctx *ctx);

/* Cross-shot stack, replaces calculate_PSDM_RTM.f90:471-497 (Cov_w = 1 on the window,
 * embedding at x_i:x_f, PSDM_sum += / Cov_sum +=; the MPI_ALLREDUCE pair :493-494 has
 * no counterpart on one GPU).  d_kernel(ny,nx_w) is added into d_psdm_sum(ny,nx_prop)
 * (FP64) at columns x_i..x_i+nx_w-1 (1-based), d_cov_sum(ny,nx_prop) (int32 fold) gains
 * 1 there, and d_kernel is cleared for the next shot.  Ordered on the context stream. */
int psdm_rtm_stack_image(psdm_rtm_ctx *ctx, float *d_kernel, int ny, int nx_w, int x_i, double *d_psdm_sum,
                         int *d_cov_sum, int nx_prop);
/* The driver's taper of the forward snapshots (calculate_PSDM_RTM.f90:421-427 with
 * taper_option != 0: taper_pad_trace, filters.f90:20-48, on every cell's Fields(j,i,:))
 * applied to the snapshots a _ram forward left in HBM under its Fields address: snapshots
 * i0 .. nt_prop2 (i0 = max(1, nt_prop2 - N_taper + 1)) are multiplied by
 * w(n) = 0.5*(1 + cos(pi*n/N_taper)), pi = 4 atan(1), in FP32 as the reference does;
 * snapshots the forward did not write stay 0 (the padding).  Bit for bit the driver's
 * result, without copying the snapshots to the host and back.  Synchronous. */
int psdm_rtm_taper_snapshots(psdm_rtm_ctx *ctx, const float *Fields, int nt_prop2, int N_taper);
/* normalize_psdm (auxiliar_PSDM.f90:1-20): d_psdm = d_psdm_sum / d_cov_sum where the
 * fold is positive, else 0. */
int psdm_rtm_normalize(psdm_rtm_ctx *ctx, const double *d_psdm_sum, const int *d_cov_sum, int ny, int nx,
                       float *d_psdm);
/* Bytes of device memory the context holds (wavefields, CPML strips, snapshots). */
long long psdm_rtm_device_bytes(const psdm_rtm_ctx *ctx);

/* Launch-shape and timing controls for the device form.
 * tile_x: columns per workgroup tile (8, 16 or 32).  use_graph: capture the time loop in
 * hipGraphs of `graph_steps` steps.  timing != 0 records HIP events around every step
 * kernel; psdm_rtm_step_stats returns their count and mean duration (ms). */
int psdm_rtm_set_launch(psdm_rtm_ctx *ctx, int tile_x, int use_graph, int graph_steps);
/* Interior columns per wavefront work item of the one-shot solver (default 2; shorter
 * items spread one shot over more wavefronts). */
int psdm_rtm_set_stream_width(psdm_rtm_ctx *ctx, int stream_w);
/* Second-derivative weights of the device solver (psdm_rtm_*_device and the engine's
 * own cfg.weights), an opt-in that has no counterpart in the reference:
 *   0 (default): the reference's form, sum over k = 0..3 of (p(i+k) + p(i-k)) * w_k
 *     (solver_space_derivs_clean.f90:1-69), bit-exact with the reference's FP32 results;
 *   1: the zero-sum form, sum over k = 1..3 of ((p(i+k) - p) + (p(i-k) - p)) * w_k.
 *     The reference's FP32 weights sum to -5.9e-8 instead of 0, a spurious mass term
 *     whose phase error grows with the record; the zero-sum form removes it (601 x 715
 *     window, 6000 steps: image 2.1e-5 rel-L2 from the FP64 form of the same algorithm
 *     instead of 1.1e-4), at the cost of no longer reproducing the reference's FP32
 *     output bit for bit.
 * Applies to every later one-shot solve on the context; the Fortran shim never sets it,
 * so the drop-in keeps the reference form. */
int psdm_rtm_set_weights(psdm_rtm_ctx *ctx, int weights);
int psdm_rtm_set_timing(psdm_rtm_ctx *ctx, int timing);
int psdm_rtm_step_stats(psdm_rtm_ctx *ctx, long long *n_steps, double *mean_ms);

/* ------------------------------------------------------------------------------
 * Batched / served engine (no counterpart in the reference, whose ranks migrate one
 * shot at a time, calculate_PSDM_RTM.f90:166-541).  `nslots` shots stay resident and
 * advance together, one launch per time step over all of them, with the steps
 * captured in hipGraphs of `chunk_steps` launches.  Shots are admitted into free slots
 * at chunk boundaries when their snapshot store fits the HBM budget, and each finished
 * shot is stacked into the engine's FP64 image (PSDM_sum) and int32 fold (Cov_sum).
 * Per shot the algorithm is the one-shot solver's, except that with store_snap > 1
 * backward step k images with the snapshot taken at forward step nt + 2 - k (the
 * reference's own bookkeeping pairs unfilled slots then, see the _ram entry points).
 * ------------------------------------------------------------------------------ */
typedef struct psdm_rtm_engine psdm_rtm_engine;

typedef struct {
    int nslots;              /* shots resident at once (1..256) */
    int ny, nx_prop;         /* migration model Vpmodel(ny, nx_prop) */
    const float *d_vel;      /* device pointer to Vpmodel, column-major */
    int max_nx, max_nt;      /* slot capacity: widest window, longest record */
    int dPML, fs, variant;   /* as the one-shot solver (variant 0 = _ram, 1 = _disk) */
    float dmodel, dt;
    int store_snap;          /* snapshot interval */
    int snap_format;         /* 0 FP32, 1 FP16, 2 BF16 (power-of-two scale per snapshot tile) */
    int kahan;               /* compensated FP32 image accumulation */
    int chunk_steps;         /* time steps per captured graph */
    int tile_x;              /* 8, 16 or 32 columns per CPML tile */
    int stream_w;            /* interior columns per wavefront work item (0 = 4) */
    int grid;                /* persistent grid size in workgroups (0 = one per work item) */
    int groups;              /* slots split into this many concurrent launch chains per captured
                                chunk, so one chain's next step fills another's drain (0 = 2; 1..4) */
    int use_graph;           /* replay captured graphs (1) or launch each step (0) */
    int max_inflight;        /* chunks queued ahead of the host (>= 1) */
    long long hbm_budget;    /* bytes for snapshot stores (0 = nslots x the largest shot) */
    double *d_psdm_sum;      /* optional caller-owned stack (ny, nx_prop) FP64, zeroed; NULL = engine-owned */
    int *d_cov_sum;          /* optional caller-owned fold (ny, nx_prop) int32, zeroed */
    int steps_per_launch;    /* time steps per launch: 0 or 1, or 2 (temporal blocking: the first step is
                                computed redundantly over a halo, so each launch reads and writes a shot's
                                wavefield once per two steps; FP32 snapshots only, ny and windows >= 16);
                                results are bit-identical to 1 */
    int weights;             /* second-derivative weights: 0 the reference's form (bit-exact with the
                                reference), 1 the zero-sum form (see psdm_rtm_set_weights); 1 needs
                                steps_per_launch <= 1 */
} psdm_rtm_engine_cfg;

typedef struct {
    int x_i, nx_w, nt;         /* window: first model column (1-based), width; record length */
    int NumRec;
    int nxSou, nySou;          /* source, 1-based inside the window */
    const int *nxRec, *nyRec;  /* host arrays, 1-based inside the window */
    const float *d_wavelet;    /* device, S(nt) */
    const float *d_data;       /* device, observed gather Data(nt, NumRec) in natural time order */
    float *d_image_out;        /* optional device (ny, nx_w): receives the shot's image */
} psdm_rtm_job;

int psdm_rtm_engine_create(psdm_rtm_ctx *ctx, const psdm_rtm_engine_cfg *cfg, psdm_rtm_engine **engine);
int psdm_rtm_engine_destroy(psdm_rtm_engine *engine);
/* Queues one shot; device buffers it names must stay valid until it completes. */
int psdm_rtm_engine_submit(psdm_rtm_engine *engine, const psdm_rtm_job *job, long long *id);
/* Issues up to max_chunks chunks (retiring and admitting shots at each boundary);
 * returns early when nothing is resident. */
int psdm_rtm_engine_run(psdm_rtm_engine *engine, int max_chunks, int *n_active, int *n_pending);
/* Shots completed since the last poll: ids and device completion times on the host
 * steady clock (seconds). */
int psdm_rtm_engine_poll(psdm_rtm_engine *engine, long long *ids, double *t_done, int cap, int *n);
/* Runs until every submitted shot is complete and stacked. */
int psdm_rtm_engine_drain(psdm_rtm_engine *engine);
/* Device pointers of the engine's stack: PSDM_sum (ny, nx_prop) FP64, Cov_sum int32. */
int psdm_rtm_engine_image(psdm_rtm_engine *engine, double **d_psdm_sum, int **d_cov_sum);
/* Mean device time of one step launch (all resident shots), from events around each
 * captured chunk of chunk_steps launches (the chunk's step-counter update included). */
int psdm_rtm_engine_timing(psdm_rtm_engine *engine, double *mean_launch_ms, long long *n_launches);
int psdm_rtm_engine_stats(psdm_rtm_engine *engine, long long *steps_issued, long long *slot_steps,
                          long long *hbm_used, long long *hbm_budget);
/* Device memory of the engine: fixed_bytes, reserved at creation for nslots shots of the
 * slot capacity (wavefields, CPML memories, images, the model's c2, descriptors, an
 * engine-owned stack); arena_bytes, the snapshot arena that admission fills shot by shot
 * (hbm_budget).  A resident shot's footprint is its slot's share of fixed_bytes plus its
 * snapshot store, so admitting by the snapshot store against the arena admits by the
 * whole footprint. */
int psdm_rtm_engine_memory(psdm_rtm_engine *engine, long long *fixed_bytes, long long *arena_bytes);

/* ------------------------------------------------------------------------------
 * Gather conditioning before migration: the per-shot steps of the driver between
 * reading a gather and migrating it (calculate_PSDM_RTM.f90:299-358), on the device.
 *   1. frequency filter when f_init > 0 and f_final > 0 (freq_filter, auxiliar.f90:
 *      498-531 -> filters_function1D, filters.f90:50-234): the 6th-order Butterworth
 *      of the reference applied to the trace zero-padded x24; typef 1 low-pass at
 *      f_final, 2 high-pass at 40 Hz, 3 band-pass f_init..f_final;
 *   2. resampling from dt_data to dt_prop when they differ (interpolation_akima.f90:
 *      1-31 -> INTRPL, akima433.f90:6-170), else the first nt_prop2 samples;
 *   3. cut_sf: samples 1..nt_sf(i) set to 0 (nt_sf != NULL);
 *   4. nt_cut == NULL: with taper_option != 0 the half-cosine taper over the N_taper
 *      samples ending at nt_prop2 (taper_pad_data, filters.f90:1-48); nt_cut != NULL
 *      (cut_multiple > 0): samples nt_cut(i)..nt_taper set to 0 and the taper ending at
 *      nt_cut(i) (taper_pad_trace);
 *   rows nt_prop2+1..nt_taper of the output are 0 (Data_target is zero-filled, :294-296).
 * d_field: device Field_Data(nt_data, NumRec) (column-major, not modified); d_target:
 * device Data_target(nt_taper, NumRec).  Steps 2-4 reproduce the reference's FP32
 * arithmetic bit for bit; the filter is the same linear operator evaluated as a
 * convolution (taps from the reference's transfer function in FP64), within 1e-5 rel-L2
 * of the reference's FFTPACK evaluation.  Synchronous on the context's stream.
 * ------------------------------------------------------------------------------ */
typedef struct {
    int nt_data;             /* samples per raw trace */
    float dt_data;           /* raw sample interval (s) */
    int typef;               /* filter type: 1 low-pass, 2 high-pass, 3 band-pass */
    float f_init, f_final;   /* filter corners (Hz); no filter unless both > 0 */
    int nt_prop2;            /* resampled samples kept */
    float dt_prop;           /* propagation time step (s) */
    int nt_taper;            /* output rows (>= nt_prop2) */
    int taper_option;        /* != 0: taper at nt_prop2 (only when nt_cut == NULL) */
    int N_taper, N_zeros;    /* taper length (>= 1 when a taper applies), zero padding */
    const int *nt_sf;        /* host (NumRec) or NULL */
    const int *nt_cut;       /* host (NumRec) or NULL; 1 <= nt_cut(i) <= nt_taper */
} psdm_rtm_cond;

int psdm_rtm_condition(psdm_rtm_ctx *ctx, const psdm_rtm_cond *cond, int NumRec, const float *d_field,
                       float *d_target);
/* The same on host arrays (the driver's Field_Data / Data_target, calculate_PSDM_RTM.f90:
 * 294-358), staged through the context's device buffers: for callers whose gathers live
 * in host memory (the Fortran driver through psdm_condition_gather). */
int psdm_rtm_condition_host(psdm_rtm_ctx *ctx, const psdm_rtm_cond *cond, int NumRec, const float *Field_Data,
                            float *Data_target);
/* With psdm_rtm_set_timing(ctx, n != 0): mean device time per call of the filter kernel
 * and of the resample/cut/taper kernel (start/stop stamps of their dispatches), and the
 * number of calls timed since timing was last set. */
int psdm_rtm_condition_timing(psdm_rtm_ctx *ctx, double *filter_ms, double *resample_ms, long long *n_calls);

/* ------------------------------------------------------------------------------
 * Depth-to-two-way-time conversion of the final image: psdm2twt (auxiliar_PSDM.f90:
 * 23-91) without its file writers.  d_vel, d_psdm: device (nz, nx) at depth step dd;
 * per column the two-way time t(k) = t(k-1) + 2*dd/v(k) (FP32, depth order), then
 * nt = 1 + ceiling(max|t| / dt) and d_vel_twt, d_psdm_twt (nt, nx) device: the velocity
 * and the image at t = (j-1)*dt by linear_interpolation_1D (linear_interpolation.f90:
 * 72-103), bit for bit.  With d_vel_twt or d_psdm_twt NULL the call only returns nt;
 * otherwise nt_cap (the outputs' rows) must be >= nt.  Each call computes the times anew.
 * Synchronous.
 * ------------------------------------------------------------------------------ */
int psdm_rtm_psdm2twt(psdm_rtm_ctx *ctx, int nx, int nz, float dd, float dt, const float *d_vel,
                      const float *d_psdm, int nt_cap, float *d_vel_twt, float *d_psdm_twt, int *nt);
/* psdm_rtm_psdm2twt that allocates nothing (no hipMalloc / hipFree, no allocating library
 * call): for callers that cannot allocate device memory (code that will be captured into a
 * HIP graph, tight serving loops, harnesses that forbid allocation).  No counterpart in the
 * reference.  Same arguments, results (bit for bit) and errors as psdm_rtm_psdm2twt; its
 * workspace -- the two-way times (nz*nx floats) and the per-column flags (nx+1 words) --
 * is a static 32 MiB device array per GPU, kept until the process ends: nz*nx + nx up to
 * about 8.3 M cells, larger sizes fail with PSDM_RTM_EINVAL.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (several threads or contexts) race.  Not
 * timed by psdm_rtm_set_timing.  Synchronous. */
int psdm_rtm_psdm2twt_noalloc(psdm_rtm_ctx *ctx, int nx, int nz, float dd, float dt, const float *d_vel,
                              const float *d_psdm, int nt_cap, float *d_vel_twt, float *d_psdm_twt, int *nt);
/* With psdm_rtm_set_timing(ctx, n != 0): mean device time per full conversion of the
 * time-accumulation kernel and of the interpolation kernel, and the conversions timed. */
int psdm_rtm_psdm2twt_timing(psdm_rtm_ctx *ctx, double *times_ms, double *interp_ms, long long *n_calls);

/* ------------------------------------------------------------------------------
 * SU trace files and shot navigation without Seismic Unix.  An SU trace is a 240-byte
 * header (60 words, size_su_header modules.f90:19) followed by ns 4-byte IEEE samples;
 * big_endian selects the byte order (endianness_data modules.f90:186, default big).
 * ------------------------------------------------------------------------------ */

/* Replaces get_shot_data (get_mcs_data.f90:150-214) and get_su_data (auxiliar.f90:9-44):
 * d_su holds ntr traces of ns samples back to back as read from the file (nbytes >=
 * ntr*(240 + 4*ns), 4-byte aligned, device); d_out (ntr, nt) device FP32, trace-major
 * (trace j's nt samples contiguous, the layout psdm_rtm_condition takes), sample i of
 * trace j = sample i - shift of the file's trace (0 outside 0..nt-1); shift 0 is the raw
 * gather, shift = int(delay_time_data/dt_mcs) the delayed one of get_mcs_data.f90:184-204
 * (whose samples the reference leaves unwritten are 0 here).  1 <= nt <= ns.  Bit for bit:
 * a pure byte reordering.  Synchronous; d_su must be complete when called. */
int psdm_rtm_su_decode(psdm_rtm_ctx *ctx, const void *d_su, long long nbytes, int ntr, int ns, int nt,
                       int big_endian, int shift, float *d_out);
/* Host-array form of psdm_rtm_su_decode for the Fortran seam (get_shot_data's own
 * arguments): reads ntr traces from byte `offset` of the file at `path` (pinned staging,
 * one host-to-device copy), decodes them on the device and copies the gather to h_out
 * (nt, ntr) column-major = (ntr, nt) trace-major.  Synchronous. */
int psdm_rtm_su_read_host(psdm_rtm_ctx *ctx, const char *path, long long offset, int ntr, int ns, int nt,
                          int big_endian, int shift, float *h_out);
/* Header words of every trace: d_out[j] = the INTEGER*4 at 1-based byte `pos` of trace j
 * (byte_fldr = 9 the shot number of get_pos_channels get_mcs_data.f90:457, byte_offset =
 * 37 the offset of :467; modules.f90:19-27).  Synchronous. */
int psdm_rtm_su_header_i32(psdm_rtm_ctx *ctx, const void *d_su, long long nbytes, int ntr, int ns, int big_endian,
                           int pos, int *d_out);
/* Replaces the SU branch of write_psdm / write_vel (auxiliar.f90:275-299, 179-210), which
 * shells out to a2b | suaddhead | sushw | suchw: d_img (ni, nj) device FP32 trace-major
 * (model(nj, ni) in the reference's column order) -> d_su, ni*(240 + 4*nj) bytes (device,
 * 4-byte aligned): tracl = tracr = trace number (from 1), ns = nj, d1 = dj, f1 = 0, d2 = di,
 * f2 = 0, cdp = int(di*(tracl-1)) as suchw computes it, trid; all other header bytes 0.
 * nj <= 65535.  Synchronous. */
int psdm_rtm_su_encode(psdm_rtm_ctx *ctx, const float *d_img, int ni, int nj, float di, float dj, int trid,
                       int big_endian, void *d_su);
/* With psdm_rtm_set_timing(ctx, n != 0): mean device time per decode / encode call. */
int psdm_rtm_su_timing(psdm_rtm_ctx *ctx, double *decode_ms, double *encode_ms, long long *n_decode,
                       long long *n_encode);

/* Host only.  Receiver and source grid indices of one MCS shot, get_pos_channels
 * (get_mcs_data.f90:418-510) and get_geometry_mcs (:128-148) in FP32: with offset_header
 * = 1 the receiver offsets come from the header words `offsets` (|offset|*offset_unit, or
 * |offset|/|offset_unit| when offset_unit < 0), with 0 from near_offset + r*drec; pos_rec
 * = pos_shot - dir_boat*offset; nxRec = 1 + floor(pos_rec/dd), nyRec = 1 +
 * floor(streamer_depth/dd), nxSou = 1 + floor(pos_shot/dd), nySou = 1 + floor(shot_depth/dd). */
int psdm_rtm_mcs_geometry(const int *offsets, int nrec, float pos_shot, int offset_header, float offset_unit,
                          int dir_boat, float near_offset, float drec, float dd, float shot_depth,
                          float streamer_depth, float *pos_rec, int *nxRec, int *nyRec, int *nxSou, int *nySou);
/* Host only.  read_nav_shot_mcs (get_mcs_data.f90:30-126): one shot per line, "shotID x"
 * (cols 11), "shotID x y" (111) or "shotID x y depth" (1111); pos = d_shot_model +
 * sqrt((x-x1)**2 + (y-y1)**2) in FP32 from the first line's (dir_boat 1) or the last
 * line's (-1) coordinates; bat = |depth|.  *shot_i / *shot_f = 0 select the first / last
 * shot (and are set to its ID); *line_i / *line_f = the 1-based lines holding them (0:
 * absent).  *n = number of lines; with cap = 0 only *n is returned. */
int psdm_rtm_nav_read(const char *path, int cols, int dir_boat, float d_shot_model, int *shot_i, int *shot_f,
                      int cap, int *ids, float *pos, float *bat, int *n, int *line_i, int *line_f);

/* ------------------------------------------------------------------------------
 * Kirchhoff depth migration of one shot (calculate_PSDM_kirchoff.f90:352-447).
 * ------------------------------------------------------------------------------ */

/* Replaces solve_eikonal_FSM (solve_eikonal_FSM.f90:1-83) and its callers
 * dynamic_window_tt_forward / _backward (dynamic_window.f90:230-325): d_vel (ny, nx)
 * device, column-major (the window Vp_w); for each of nsrc source points (src_x, src_z:
 * host, 1-based column and row in the window) the travel-time field, d_tt (ny, nx, nsrc)
 * device, column-major = TT_r(ny, nx_w, NumRec) (or TT_s for one source).  Four fast
 * sweeps from T = 1e9, T(source) = 0, bit for bit.  nx <= 8192.  Synchronous.
 * Device workspace held by the ctx: 4 * (2 * (nx + ny - 1) * nx + 1024) bytes per source
 * solved at once plus 8 * (nx + ny - 1) * nx for 1/v.  The sources run in batches of as
 * many as the workspace limit holds (psdm_rtm_set_eikonal_workspace, default 16 GiB: all
 * 481 of a 601 x 715 window, 71 of an 801 x 5102 one), at least one per batch. */
int psdm_rtm_eikonal(psdm_rtm_ctx *ctx, const float *d_vel, int ny, int nx, float dx, float dy, int nsrc,
                     const int *src_x, const int *src_z, float *d_tt);
/* Host-array form of psdm_rtm_eikonal for the Fortran seam (dynamic_window_tt_forward /
 * _backward's arrays): h_vel (ny, nx) and h_tt (ny, nx, nsrc) column-major host arrays.
 * The device holds one batch of travel times at a time.  Synchronous. */
int psdm_rtm_eikonal_host(psdm_rtm_ctx *ctx, const float *h_vel, int ny, int nx, float dx, float dy, int nsrc,
                          const int *src_x, const int *src_z, float *h_tt);
/* No counterpart in the reference (its solver keeps one source's arrays at a time): the
 * limit on the eikonal workspace of this ctx, in bytes (> 0; default 16 GiB). */
int psdm_rtm_set_eikonal_workspace(psdm_rtm_ctx *ctx, size_t max_bytes);
/* The eikonal's quotients by dx*dx, dy*dy and 2*aa (local_solver, solve_eikonal_FSM.f90:63-79)
 * are correctly rounded either way; bit i of *fast_mask (i = 0, 1, 2 in that order) is set when
 * that divisor uses a product and an FMA in place of the IEEE division, having matched it
 * for every significand on the device (checked on the first use of each divisor value and
 * kept by the ctx; psdm_rtm_kirchhoff_image does the same for its divisions by dt and
 * 2*dt).  No counterpart in the reference. */
int psdm_rtm_eikonal_division(psdm_rtm_ctx *ctx, float dx, float dy, int *fast_mask);
/* *nb = the sources psdm_rtm_eikonal solves per launch for this shape under the ctx's
 * workspace limit (1 <= *nb <= nsrc). */
int psdm_rtm_eikonal_batch(psdm_rtm_ctx *ctx, int ny, int nx, int nsrc, int *nb);
/* The imaging loop of calculate_PSDM_kirchoff.f90:394-447: d_img (ny, nx) device =
 * sum over receivers k (in order) of weight * (d(it+1) - d(it-1)) / (2*dt) with
 * it = 1 + floor((TT_s + TT_r(k))/dt) (skipped unless 1 < it < nt_prop), Veff = vp0,
 * weight = w_theta * min(cos_theta / sqrt(rs*rr + 1e-6), wmax), the aperture taper
 * w_theta = exp(-((theta - theta_max)/theta_taper)**2) beyond theta_max.  d_tts (ny, nx),
 * d_ttr (ny, nx, nrec) from psdm_rtm_eikonal; d_data (nrec, ld) device, trace-major
 * (Data_target(nt_taper, NumRec) in the reference's order), ld >= nt_prop.  The travel-time
 * index is exact; acos and exp are the device's (within a few ulp of the host's libm).
 * Synchronous. */
int psdm_rtm_kirchhoff_image(psdm_rtm_ctx *ctx, const float *d_tts, const float *d_ttr, const float *d_data, int ld,
                             int nrec, int nt_prop, int ny, int nx, float dd, float dt, float vp0, float theta_max,
                             float theta_taper, float wmax, float *d_img);
/* With psdm_rtm_set_timing(ctx, n != 0): mean device time per eikonal / imaging call. */
int psdm_rtm_kirchhoff_timing(psdm_rtm_ctx *ctx, double *eikonal_ms, double *image_ms, long long *n_eikonal,
                              long long *n_image);

#ifdef __cplusplus
}
#endif
#endif /* PSDM_RTM_H */
