// Internal types shared by the RTM kernels and the host runtime (libpsdm_rtm.so).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace psdm {

// Central-difference weights of accuracy 3 (radius 3), finite_difference_coef
// (solver_FD_coeffs.f90:14-19 first derivative, :43-50 and :68 second derivative
// with the centre weight halved).  Written as the reference's FP32 quotients.
constexpr float kC1_1 = 3.f / 4.f, kC1_2 = -3.f / 20.f, kC1_3 = 1.f / 60.f;
constexpr float kC2_0 = (-49.f / 18.f) / 2.f, kC2_1 = 3.f / 2.f, kC2_2 = -3.f / 20.f, kC2_3 = 1.f / 90.f;

constexpr int kTileY = 64;        // rows per tile: one wavefront, one row per lane
constexpr int kBlock = 256;       // 4 wavefronts per workgroup
constexpr int kMaxTables = 1024;  // max CPML cells per side (dPML + 6)

// Snapshot storage formats.  FP16/BF16 snapshots carry one power-of-two scale per
// (snapshot, tile): the tile's max-abs value is mapped to 2^14 (FP16) or 1 (BF16), so
// the only error is the 16-bit rounding of the scaled value.
enum SnapFmt { kSnapF32 = 0, kSnapF16 = 1, kSnapBF16 = 2 };

// Everything one time step of one shot needs.  Layout in HBM:
//   wavefields  p[i*ld + j], i in [0,nxt), j in [0,nyt), ld = nyt rounded up to 64
//   x-CPML mem  x[c*ld + j], c in [0,2dn): c < dn left strip column c, else right
//               strip column c-dn (grid column nx+c); all nyt rows (strips+corners)
//   y-CPML mem  y[i*2dn + r], r in [0,2dn): r < dn top row r, else bottom row ny+r;
//               all nxt columns (strips+corners)
//   c2          the unpadded window c2w[(i-dn)*ldc + (j-dn)], read with clamped
//               indices: the reference replicates edge velocities into the pads
//   snapshots   F[s][i][j] over the interior (ny x nx), s = snapshot slot
// Region-aligned tiling: tile rows cover the top CPML [0,dn), the interior [dn,dn+ny)
// and the bottom CPML separately (64 rows each, the last one of a region partial);
// tile columns likewise with TX columns.  Every tile is then wholly interior, x strip,
// y strip or corner.  Tiles are numbered column-major: t = C * nrt + R.
struct Tiling {
    int rt_pml, rt_int, ct_pml, ct_int, nrt, nct, ntiles;
};

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
// Leading dimension of the interior arrays (snapshots, image, c2): whole 64-float
// column segments, so a wavefront's 64 rows are one aligned 256-byte access.
__host__ __device__ inline int interior_ld(int ny) { return cdiv(ny, 64) * 64; }

__host__ __device__ inline Tiling make_tiling(int ny, int nx, int dn, int tx) {
    Tiling T;
    T.rt_pml = cdiv(dn, kTileY);
    T.rt_int = cdiv(ny, kTileY);
    T.ct_pml = cdiv(dn, tx);
    T.ct_int = cdiv(nx, tx);
    T.nrt = 2 * T.rt_pml + T.rt_int;
    T.nct = 2 * T.ct_pml + T.ct_int;
    T.ntiles = T.nrt * T.nct;
    return T;
}

// Region (0 interior, 1 low-side CPML, 2 high-side CPML) and [lo, hi) of tile index r
// along one axis with `npml` CPML tiles on each side sharing its `n_pml` cells evenly
// (so every CPML tile spans at least 4 cells when n_pml >= 7 and w >= 8: the clamped
// edge stencils then stay inside the tile's staged halo), and `nint` interior tiles of
// `w` cells over `n_int` cells.
__host__ __device__ inline int axis_tile(int r, int npml, int nint, int n_pml, int n_int, int w, int *lo, int *hi) {
    if (r < npml) {
        *lo = r * n_pml / npml;
        *hi = (r + 1) * n_pml / npml;
        return 1;
    }
    r -= npml;
    if (r < nint) {
        *lo = n_pml + r * w;
        *hi = *lo + w < n_pml + n_int ? *lo + w : n_pml + n_int;
        return 0;
    }
    r -= nint;
    *lo = n_pml + n_int + r * n_pml / npml;
    *hi = n_pml + n_int + (r + 1) * n_pml / npml;
    return 2;
}

// Tile index along one axis of cell c (inverse of axis_tile).
__host__ __device__ inline int axis_tile_of(int c, int npml, int n_pml, int n_int, int w) {
    if (c >= n_pml && c < n_pml + n_int) return npml + (c - n_pml) / w;
    const int side = c < n_pml ? 0 : 1;
    const int cl = side ? c - n_pml - n_int : c;
    int r = 0;
    while ((r + 1) * n_pml / npml <= cl) ++r;
    return side ? npml + cdiv(n_int, w) + r : r;
}

// Tile index of padded cell (j, i).
__host__ __device__ inline int tile_of_cell(const Tiling &T, int j, int i, int ny, int nx, int dn, int tx) {
    const int R = axis_tile_of(j, T.rt_pml, dn, ny, kTileY);
    const int C = axis_tile_of(i, T.ct_pml, dn, nx, tx);
    return C * T.nrt + R;
}

struct StepArgs {
    int ny, nx, dn, nyt, nxt, ld;
    int ntile_y, ntiles, per_xcd;  // ntile_y = Tiling::nrt
    int rt_pml, rt_int, ct_pml, ct_int;
    int fs, ram;      // free surface flag; ram = 1 -> _ram semantics, 0 -> _disk
    int k;            // time step index of the reference loop, 2..nt
    int mode;         // 0 forward (snapshots), 1 backward (imaging)
    int snap_slot;    // snapshot slot this step writes/reads, -1 = none
    int zc, zp, zm;   // state not yet written this phase reads as zero: p(k-1), p(k-2), CPML memory
    int snap_fmt;     // SnapFmt
    int kahan;        // compensated image accumulation
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
    const int *ev_ptr;     // per tile [ev_ptr[t], ev_ptr[t+1]) injection events
    const int *ev_cell;    // event cell j + i*nyt
    const int *ev_inj;     // column of inj_val
    const float *inj_val;  // (nt-1) x ninj, row k-2 holds the values injected at step k
    int ninj;
    void *snaps;           // forward: snapshot base (format snap_fmt); backward: read
    float *snap_scale;     // FP16/BF16: scale[slot * ntiles + tile]
    float *image;          // backward: kernel(ny,nx), leading dimension lds
    float *image_comp;     // Kahan compensation (kahan != 0)
    int lds;               // leading dimension of snapshots, image and c2: ny rounded up to 64
    size_t snap_stride;    // elements per snapshot (lds*nx)
    const int *rec_cell;   // receivers: cell j + i*ld, or -1 outside the grid
    float *rec;            // Rec_point(nt, NumRec) column-major, leading dimension rec_ld = nt
    int rec_ld;
    int nrec;
    int rec_row;           // row of rec that p_c (step k-1) fills, -1 = none
};

// One slot of the batched engine: a shot resident on the device for its forward and
// backward phases.  `a` holds the static fields; the per-step pointers and flags are
// resolved on the device from the global step (rtm_batch_kernel).
struct SlotDesc {
    StepArgs a;
    float *pA, *pB, *phxA, *phxB, *phyA, *phyB, *hsA, *hsB;
    long long g0;  // global step of the slot's first forward step (k = 2); < 0: slot idle
    int nt, ss, nsnap;
    const int *ev_ptr_f, *ev_cell_f, *ev_inj_f;
    const float *inj_f;
    int ninj_f;
    const int *ev_ptr_b, *ev_cell_b, *ev_inj_b;
    const float *inj_b;
    int ninj_b;
};

struct BatchArgs {
    const SlotDesc *slots;
    int nslots;
    int cap_tiles;           // work items per slot (tiles of the slot capacity)
    int nwork;               // nslots * cap_tiles
    int per_xcd;             // work items per XCD
    const long long *gbase;  // global step of the chunk's first step
    int goff;                // this launch's offset inside the chunk
};

void launch_step(const StepArgs &a, int tile_x, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_batch(const BatchArgs &b, int tile_x, int grid, hipStream_t s, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
void launch_advance(long long *gbase, long long by, hipStream_t s);
void launch_record(const float *p, const int *rec_cell, float *rec, int rec_ld, int row, int nrec, hipStream_t s);
void launch_c2(const float *vel, float *c2, int ny, int nx, int ldc, float dt, float dx, hipStream_t s);
void launch_stack(float *img, float *comp, int ld_img, int ny, int nx_w, int x0, double *sum, int ld_sum, int *fold,
                  hipStream_t s);
void launch_normalize(const double *sum, const int *fold, float *out, size_t n, hipStream_t s);
// out[r*ncols + e] = traces[cols[e]*nt + (reverse ? nt-1-r : r)], r in [0, nt-1)
void launch_gather(const float *traces, int nt, const int *cols, int ncols, float *out, int reverse, hipStream_t s);

}  // namespace psdm
