// Internal types shared by the RTM kernels and the host runtime (libpsdm_rtm.so).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fastdiv.hpp"

namespace psdm {

// Central-difference weights of accuracy 3 (radius 3), finite_difference_coef
// (solver_FD_coeffs.f90:14-19 first derivative, :43-50 and :68 second derivative
// with the centre weight halved).  Written as the reference's FP32 quotients.
constexpr float kC1_1 = 3.f / 4.f, kC1_2 = -3.f / 20.f, kC1_3 = 1.f / 60.f;
constexpr float kC2_0 = (-49.f / 18.f) / 2.f, kC2_1 = 3.f / 2.f, kC2_2 = -3.f / 20.f, kC2_3 = 1.f / 90.f;

constexpr int kTileY = 64;        // CPML tiles: rows per tile, one row per lane
constexpr int kBlock = 256;       // 4 wavefronts per workgroup
constexpr int kMaxTables = 1024;  // max CPML cells per side (dPML + 6)
constexpr int kStreamRows = 248;  // interior stream items: rows owned per wavefront
constexpr int kStageCols = 4;     // FP32 snapshot columns a stream item stages in LDS
constexpr int kIntPad = 4;        // front rows of padding of the interior arrays

// Snapshot storage formats.  FP16/BF16 snapshots carry one power-of-two scale per
// (snapshot, interior column, 248-row stream tile): the max-abs value there is mapped
// to 2^14 (FP16) or 1 (BF16), so the only error is the 16-bit rounding of the scaled
// value.
enum SnapFmt { kSnapF32 = 0, kSnapF16 = 1, kSnapBF16 = 2 };
// Snapshot / imaging variant a step kernel is compiled for (the interior carries only the
// registers its variant needs): FP32 or 16-bit snapshots, plain or Kahan imaging.
enum SnapVariant { kSvF32 = 0, kSvF32Kahan = 1, kSv16 = 2, kSv16Kahan = 3 };
// Added to a SnapVariant: the kernel's second derivatives use the zero-sum weights (d2z).
constexpr int kSvZeroSum = 4;
constexpr int snap_variant(int snap_fmt, int kahan) { return (snap_fmt != kSnapF32 ? 2 : 0) + (kahan ? 1 : 0); }

// Everything one time step of one shot needs.  Layout in HBM:
//   wavefields  p[i*ld + j], i in [0,nxt), j in [0,nyt): the pointer is offset into its
//               buffer (state_off) so that row dn-4 of every column is 128-byte aligned
//   x-CPML mem  x[c*ld + j], c in [0,2dn): c < dn left strip column c, else right
//               strip column c-dn (grid column nx+c); all nyt rows (strips+corners)
//   y-CPML mem  y[i*2dn + r], r in [0,2dn): r < dn top row r, else bottom row ny+r;
//               all nxt columns (strips+corners)
//   c2          the unpadded window c2w[(i-dn)*ldc + (j-dn)], read with clamped
//               indices by the CPML cells (the reference replicates edge velocities
//               into the pads); the pointer is kIntPad rows into its buffer
//   snapshots   F[s][(i-dn)*lds + (j-dn)] over the interior (ny x nx), s = snapshot slot;
//   image       same layout (kIntPad front rows in the buffer, like c2)
// Work of one step:
//   CPML tiles  64 rows x TX columns of the region-aligned grid: tile rows cover the
//               top CPML [0,dn), the interior [dn,dn+ny) and the bottom CPML apart,
//               columns likewise, so each tile is x strip, y strip or corner.  Only
//               the non-interior tiles of that grid are used: first the top/bottom
//               strips (tb_rows row tiles x the listed column tiles, c = C*tb_rows + r;
//               under the _ram rule the interior column tiles from column dn on are
//               left out, ctb_skip of them, and with a free surface the top row tiles),
//               then the x strips over the interior rows
//               (2*ct_pml x rt_int, c = cc*rt_int + r).
//   stream items  the interior in wavefront-sized pieces: kStreamRows rows (4 per lane,
//               one 4-row halo lane at each end) x sw columns, walked column by column
//               with a 7-column register window (ri row tiles x cs column segments,
//               u = CS*ri + R); four items per workgroup.
struct Tiling {
    int rt_pml, rt_int, ct_pml, ct_int, nrt, nct;  // the region-aligned 64 x TX grid
    int ncpml_a, ncpml;                            // CPML tiles: top/bottom strips, all
    int ctb_skip;                                  // interior column tiles left out of the top/bottom strips
    int tb_rows;                                   // row tiles listed per top/bottom column (bottom only with fs)
    int ri, cs, sw, nstream, nsblk;                // stream items and their workgroups
    int nitems;                                    // workgroups of one step: ncpml + nsblk
};

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
// Leading dimension of the interior arrays (snapshots, image, c2): a multiple of 64
// holding the kIntPad front rows and every row a stream item reads (its halo lanes).
__host__ __device__ inline int interior_ld(int ny) { return cdiv((ny > kStreamRows ? ny : kStreamRows) + 8, 64) * 64; }
// Row offset of the wavefield pointer into its buffer: row dn-4 lands on a 32-float
// (128-byte) boundary, so every stream item's 256-row column reads are 16-byte aligned.
__host__ __device__ inline int state_off(int dn) { return ((4 - dn) % 32 + 32) % 32; }
// Leading dimension of the wavefield arrays: holds state_off + the nyt grid rows and
// every row a stream item reads.
__host__ __device__ inline int state_ld(int ny, int dn) {
    const int rows = ny + 2 * dn > dn + (ny > kStreamRows ? ny : kStreamRows) + 8
                         ? ny + 2 * dn
                         : dn + (ny > kStreamRows ? ny : kStreamRows) + 8;
    return cdiv(state_off(dn) + rows, 64) * 64;
}
// First interior row (j - dn) a stream item of row tile R reads at lane 1: 248 R, except
// that the last tile is moved up (to a multiple of 4) so it ends at or below ny + 3.
__host__ __device__ inline int stream_origin(int R, int ny) {
    if (ny <= kStreamRows) return 0;
    const int last = cdiv(ny - kStreamRows, 4) * 4;
    return R * kStreamRows < last ? R * kStreamRows : last;
}

// ram: the _ram padding rule, under which the top and bottom pads past interior column dn
// have c2 = 0 and p stays exactly 0 (solver_ac_forward_ram.f90:288-302); their tiles are
// left out of the work list, and the top/bottom strip tiles list the corner columns and
// the interior column tiles that start before column dn only.  fs: with a free surface
// the top strip and its corners are never updated (:391, :400-416) and p stays 0 there,
// so only the bottom row tiles are listed.  Either way the caller zeroes both wavefield
// buffers before a shot's first step.
__host__ __device__ inline Tiling make_tiling(int ny, int nx, int dn, int tx, int sw, bool ram, bool fs) {
    Tiling T;
    T.rt_pml = cdiv(dn, kTileY);
    T.rt_int = cdiv(ny, kTileY);
    T.ct_pml = cdiv(dn, tx);
    T.ct_int = cdiv(nx, tx);
    T.nrt = 2 * T.rt_pml + T.rt_int;
    T.nct = 2 * T.ct_pml + T.ct_int;
    const int live = ram ? (T.ct_int < cdiv(dn, tx) ? T.ct_int : cdiv(dn, tx)) : T.ct_int;
    T.ctb_skip = T.ct_int - live;
    T.tb_rows = (fs ? 1 : 2) * T.rt_pml;
    T.ncpml_a = T.tb_rows * (T.nct - T.ctb_skip);
    T.ncpml = T.ncpml_a + 2 * T.ct_pml * T.rt_int;
    T.ri = cdiv(ny, kStreamRows);
    T.sw = sw;
    T.cs = cdiv(nx, sw);
    T.nstream = T.ri * T.cs;
    T.nsblk = cdiv(T.nstream, 4);
    T.nitems = T.ncpml + T.nsblk;
    return T;
}

// Region (0 interior, 1 low-side CPML, 2 high-side CPML) and [lo, hi) of tile index r
// along one axis with `npml` CPML tiles on each side sharing its `n_pml` cells evenly
// (so every CPML tile spans at least 4 cells when n_pml >= 7 and w >= 8: the clamped
// edge stencils then stay inside the tile's staged halo), and `nint` interior tiles of
// `w` cells over `n_int` cells.  `fpml` divides by npml.
__host__ __device__ inline int axis_tile(int r, int npml, int nint, int n_pml, int n_int, int w, const FastDiv &fpml,
                                         int *lo, int *hi) {
    if (r < npml) {
        *lo = fdiv(r * n_pml, fpml);
        *hi = fdiv((r + 1) * n_pml, fpml);
        return 1;
    }
    r -= npml;
    if (r < nint) {
        *lo = n_pml + r * w;
        *hi = *lo + w < n_pml + n_int ? *lo + w : n_pml + n_int;
        return 0;
    }
    r -= nint;
    *lo = n_pml + n_int + fdiv(r * n_pml, fpml);
    *hi = n_pml + n_int + fdiv((r + 1) * n_pml, fpml);
    return 2;
}

// Stream item that owns interior cell (jw, iw) (0-based inside the window).
__host__ __device__ inline int stream_item_of_cell(const Tiling &T, int jw, int iw) {
    return (iw / T.sw) * T.ri + jw / kStreamRows;
}

struct StepArgs {
    int ny, nx, dn, nyt, nxt, ld;
    int nitems, per_xcd;  // workgroups of one step (Tiling::nitems); per XCD
    int rt_pml, rt_int, ct_pml, ct_int;
    int ncpml_a, ncpml, ri, sw, nstream;
    int ctb_skip;  // Tiling::ctb_skip
    int tb_rows;   // Tiling::tb_rows
    FastDiv fd_tbr, fd_rint, fd_ri, fd_rpml, fd_cpml;  // division by tb_rows, rt_int, ri, rt_pml, ct_pml
    int fs, ram;      // free surface flag; ram = 1 -> _ram semantics, 0 -> _disk
    int k;            // time step index of the reference loop, 2..nt
    int mode;         // 0 forward (snapshots), 1 backward (imaging)
    int snap_slot;    // snapshot slot this step writes/reads, -1 = none
    int zc, zp, zm;   // state not yet written this phase reads as zero: p(k-1), p(k-2), CPML memory
    int snap_fmt;     // SnapFmt
    int kahan;        // compensated image accumulation
    int zsum;         // host side: launch the zero-sum-weight kernels (d2z) instead of the reference's (d2)
    float dt, ss_f;   // dt; float(store_snap)
    const float *pc;  // p at step k-1
    float *pn;        // in: p at step k-2; out: p at step k
    const float *phx_o;
    float *phx_n, *chx;
    const float *phy_o;
    float *phy_n, *chy;
    const float *hs_c;  // _ram with fs: own history of the x strips' first interior row
    float *hs_n;        // (2dn values), in: k-2, out: k
    const float *c2;
    int ldc;
    const float *fA, *fB;  // 1D damping tables (dn), index = distance into the CPML - 1
    const int *ev_ptr;     // per stream item [ev_ptr[u], ev_ptr[u+1]) injection events, by column
    const int *ev_cell;    // event cell j + i*nyt
    const int *ev_inj;     // column of inj_val (the one-step lists are built with ev_inj[e] == e,
                           // which stream_item relies on: it reads event e's value in column e)
    const float *inj_val;  // (nt-1) x ninj, row k-2 holds the values injected at step k
    int ninj;
    void *snaps;           // forward: snapshot base (format snap_fmt); backward: read
    float *snap_scale;     // FP16/BF16: scale[slot * scale_stride + R * nx + (i - dn)]
    int scale_stride;      // ri * nx
    float *image;          // backward: kernel(ny,nx), leading dimension lds
    float *image_comp;     // Kahan compensation (kahan != 0)
    int lds;               // leading dimension of snapshots and image (interior_ld)
    size_t snap_stride;    // elements per snapshot (lds*nx)
    const int *rec_cell;   // receivers: cell j + i*ld, or -1 outside the grid
    float *rec;            // Rec_point(nt, NumRec) column-major, leading dimension rec_ld = nt
    int rec_ld;
    int nrec;
    int rec_row;           // row of rec that p_c (step k-1) fills, -1 = none
};

// Copies a tiling into the step arguments (host side).
inline void set_tiling(StepArgs &a, const Tiling &T) {
    a.nitems = T.nitems;
    a.rt_pml = T.rt_pml;
    a.rt_int = T.rt_int;
    a.ct_pml = T.ct_pml;
    a.ct_int = T.ct_int;
    a.ncpml_a = T.ncpml_a;
    a.ctb_skip = T.ctb_skip;
    a.tb_rows = T.tb_rows;
    a.ncpml = T.ncpml;
    a.ri = T.ri;
    a.sw = T.sw;
    a.nstream = T.nstream;
    a.fd_tbr = make_fastdiv(T.tb_rows);
    a.fd_rint = make_fastdiv(T.rt_int);
    a.fd_ri = make_fastdiv(T.ri);
    a.fd_rpml = make_fastdiv(T.rt_pml);
    a.fd_cpml = make_fastdiv(T.ct_pml);
}

// Two time steps per launch (temporal blocking, engine only).  One launch advances a shot
// from levels (k-1, k-2) to (k+1, k) with no dependency between its work items: every item
// computes the first step redundantly over a halo wide enough for the second.
//   core items   the interior rows [kTbTop, ny-3) and columns [3, nx-3): wavefront stream
//                items of 4 rows per lane whose lanes 2..61 own kTbRows rows (lanes 0-1 and
//                62-63 hold the halo), walking sw columns with the first step 3 columns
//                ahead of the second; their halo never leaves the interior;
//   edge tiles   everything else -- the CPML strips and a band of 3 (kTbTop on top) interior
//                cells along them -- in tiles of at most kTbTY rows x kTbTX columns, each
//                computing the first step on its tile +-3 (+-6 for the CPML memory) in LDS.
// Regions of the edge (padded grid rows x columns): top [0, dn+kTbTop) x all, bottom
// [dn+ny-3, nyt) x all, left [dn+kTbTop, dn+ny-3) x [0, dn+3), right (same rows) x
// [dn+nx-3, nxt).  Each region splits evenly into ceil(h/kTbTY) x ceil(w/kTbTX) tiles.
constexpr int kTbTop = 4;    // interior rows above the core (the core's first row is 16-byte aligned)
constexpr int kTbBand = 3;   // interior columns left/right of the core and rows below it
constexpr int kTbRows = 240; // core rows owned per wavefront
constexpr int kTbTY = 52, kTbTX = 16;  // edge tile: at most 52 rows x 16 columns owned
constexpr int kTbH = 9;      // edge tile LDS halo: the first step's memory needs +-6, its stencil +-3

struct TbTiling {
    int er_top, er_mid, er_bot;  // edge row tiles of the top, middle (left/right) and bottom regions
    int ec_all, ec_side;         // edge column tiles across the full width, and of a side region
    int ne;                      // edge tiles: (er_top + er_bot) * ec_all + 2 * er_mid * ec_side
    int cr, cs, sw, ncore, ncblk;  // core row tiles, column segments, width, items, workgroups
    int nitems;                  // workgroups of one launch: ne + ncblk
};

// First interior row read by lane 0 + 8 of core row tile R: kTbTop + kTbRows * R, the last
// tile moved up (to a multiple of 4) so it ends at or below ny - 3.
__host__ __device__ inline int tb_origin(int R, int ny) {
    const int o = kTbTop + kTbRows * R, last = cdiv(ny - 3 - kTbRows, 4) * 4;
    return o < last || last < kTbTop ? o : last;
}

__host__ __device__ inline TbTiling make_tb_tiling(int ny, int nx, int dn, int sw) {
    TbTiling T;
    T.er_top = cdiv(dn + kTbTop, kTbTY);
    T.er_bot = cdiv(dn + 3, kTbTY);
    T.er_mid = cdiv(ny - 3 - kTbTop, kTbTY);
    T.ec_all = cdiv(nx + 2 * dn, kTbTX);
    T.ec_side = cdiv(dn + 3, kTbTX);
    T.ne = (T.er_top + T.er_bot) * T.ec_all + 2 * T.er_mid * T.ec_side;
    T.cr = cdiv(ny - 3 - kTbTop, kTbRows);
    T.sw = sw;
    T.cs = cdiv(nx - 6, sw);
    T.ncore = T.cr * T.cs;
    T.ncblk = cdiv(T.ncore, 4);
    T.nitems = T.ne + T.ncblk;
    return T;
}

// [lo, hi) of the r-th of n even parts of [a, b).
__host__ __device__ inline void tb_split(int r, int n, int a, int b, int *lo, int *hi) {
    *lo = a + (int)((long long)r * (b - a) / n);
    *hi = a + (int)((long long)(r + 1) * (b - a) / n);
}

// Owned rectangle [j0, j1) x [i0, i1) of edge tile e (padded grid).
__host__ __device__ inline void tb_edge_tile(const TbTiling &T, int e, int ny, int nx, int dn, int *j0, int *j1,
                                             int *i0, int *i1) {
    const int nyt = ny + 2 * dn, nxt = nx + 2 * dn;
    const int ntop = T.er_top * T.ec_all, nbot = T.er_bot * T.ec_all, nside = T.er_mid * T.ec_side;
    if (e < ntop) {
        tb_split(e / T.ec_all, T.er_top, 0, dn + kTbTop, j0, j1);
        tb_split(e % T.ec_all, T.ec_all, 0, nxt, i0, i1);
    } else if ((e -= ntop) < nbot) {
        tb_split(e / T.ec_all, T.er_bot, dn + ny - 3, nyt, j0, j1);
        tb_split(e % T.ec_all, T.ec_all, 0, nxt, i0, i1);
    } else {
        e -= nbot;
        const int right = e >= nside;
        if (right) e -= nside;
        tb_split(e / T.ec_side, T.er_mid, dn + kTbTop, dn + ny - 3, j0, j1);
        if (right)
            tb_split(e % T.ec_side, T.ec_side, dn + nx - 3, nxt, i0, i1);
        else
            tb_split(e % T.ec_side, T.ec_side, 0, dn + 3, i0, i1);
    }
}

// One slot of the batched engine: a shot resident on the device for its forward and
// backward phases.  `a` holds the static fields; the per-step pointers and flags are
// resolved on the device from the global step (rtm_batch_kernel).
struct SlotDesc {
    StepArgs a;
    float *pA, *pB, *phxA, *phxB, *phyA, *phyB, *hsA, *hsB;
    long long g0;  // global step of the slot's first forward step (k = 2); < 0: slot idle
    int nt, ss, nsnap;
    // the slot's range [woff, woff + witems) in its launch group's compact work list (the
    // work items of the group's occupied slots back to back; witems = 0 for an idle slot).
    // Set in a group's first slot only: wtotal, the list's length; wuni, the items of every
    // slot when all of the group's slots hold that many (> 0), else 0; wfd, division by
    // wuni, or else by the mean items per slot (the first guess of an item's slot)
    int woff, witems, wtotal, wuni;
    FastDiv wfd;
    FastDiv fd_ss;  // division by ss
    const int *ev_ptr_f, *ev_cell_f, *ev_inj_f;
    const float *inj_f;
    int ninj_f;
    const int *ev_ptr_b, *ev_cell_b, *ev_inj_b;
    const float *inj_b;
    int ninj_b;
    // two steps per launch: the second pair of wavefield levels (pA/pB hold levels
    // (k-1, k-2) of even launches of a phase, pC/pD those of odd ones), second buffers of
    // the chi memories and the x-strip history, the two-step tiling, and the injection
    // events per two-step work item (the cells each item computes, halo included)
    float *pC, *pD, *chxB, *chyB, *hsC, *hsD;
    TbTiling tb;
    // the core kernel's compact work list (as woff ... wfd, over tb.ncblk workgroups per slot);
    // woff ... wfd then hold the edge kernel's (tb.ne tiles per slot)
    int cwoff, cwtotal, cwuni;
    FastDiv cwfd;
    const int *tev_ptr_f, *tev_cell_f, *tev_inj_f;
    const int *tev_ptr_b, *tev_cell_b, *tev_inj_b;
};

// One launch group: slots[0 .. nslots) (their SlotDesc woff / witems give the group's
// compact work list); the grid is sized on the host for every slot at capacity.
struct BatchArgs {
    const SlotDesc *slots;
    int nslots;
    const long long *gbase;  // global step of the chunk's first step
    int goff;                // this launch's offset inside the chunk
    // compact = 0: slot s owns items [s*cap_tiles, (s+1)*cap_tiles) (the work items of the
    // slot capacity), exact when every slot holds a full-capacity window and read from the
    // launch arguments alone; compact = 1: the group's compact work list (SlotDesc woff)
    int compact;
    int cap_tiles;
    FastDiv fd_cap;  // division by cap_tiles
};

void launch_step(const StepArgs &a, int tile_x, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// sv: the SnapVariant of every slot (the engine's snapshot format and Kahan setting)
void launch_batch(const BatchArgs &b, int tile_x, int sv, int grid, hipStream_t s, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
void launch_advance(long long *gbase, long long by, hipStream_t s);
// Two steps per launch (rtm_tblock.hip): part 0 the edge tiles (list woff..., cap_tiles =
// edge tiles per slot at capacity), part 1 the core items (list cwoff..., cap_tiles = core
// workgroups per slot); sv: kSvF32 or kSvF32Kahan.
void launch_tb_batch(const BatchArgs &b, int part, int sv, int grid, hipStream_t s, hipEvent_t e0 = nullptr,
                     hipEvent_t e1 = nullptr);
void launch_record(const float *p, const int *rec_cell, float *rec, int rec_ld, int row, int nrec, hipStream_t s);
void launch_c2(const float *vel, float *c2, int ny, int nx, int ldc, float dt, float dx, hipStream_t s);
void launch_stack(float *img, float *comp, int ld_img, int ny, int nx_w, int x0, double *sum, int ld_sum, int *fold,
                  hipStream_t s);
void launch_normalize(const double *sum, const int *fold, float *out, size_t n, hipStream_t s);
// snaps[(s0+q)*stride + i*lds + j] *= w[q], q < nq, i < nx, j < ny
void launch_taper_snaps(float *snaps, size_t stride, int lds, int ny, int nx, int s0, int nq, const float *w,
                        hipStream_t s);
// out[r*ncols + e] = traces[cols[e]*nt + (reverse ? nt-1-r : r)], r in [0, nt-1)
void launch_gather(const float *traces, int nt, const int *cols, int ncols, float *out, int reverse, hipStream_t s);

}  // namespace psdm
