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

// Everything one time step of one shot needs.  Layout in HBM:
//   wavefields  p[i*ld + j], i in [0,nxt), j in [0,nyt), ld = nyt rounded up to 64
//   x-CPML mem  x[c*ld + j], c in [0,2dn): c < dn left strip column c, else right
//               strip column c-dn (grid column nx+c); all nyt rows (strips+corners)
//   y-CPML mem  y[i*2dn + r], r in [0,2dn): r < dn top row r, else bottom row ny+r;
//               all nxt columns (strips+corners)
//   c2          the unpadded window c2w[(i-dn)*ldc + (j-dn)], read with clamped
//               indices: the reference replicates edge velocities into the pads
//   snapshots   F[s][i][j] over the interior (ny x nx), s = snapshot slot
struct StepArgs {
    int ny, nx, dn, nyt, nxt, ld;
    int ntile_y, ntiles, per_xcd;
    int fs, ram;      // free surface flag; ram = 1 -> _ram semantics, 0 -> _disk
    int k;            // time step index of the reference loop, 2..nt
    int mode;         // 0 forward (snapshots), 1 backward (imaging)
    int snap_slot;    // snapshot slot this step writes/reads, -1 = none
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
    float *snaps;          // forward: snapshot base; backward: read
    float *image;          // backward: kernel(ny,nx), ld = ny
    size_t snap_stride;    // floats per snapshot (ny*nx)
    const int *rec_cell;   // receivers: cell j + i*ld, or -1 outside the grid
    float *rec;            // Rec_point(nt, NumRec) column-major, leading dimension rec_ld = nt
    int rec_ld;
    int nrec;
    int rec_row;           // row of rec that p_c (step k-1) fills, -1 = none
};

void launch_step(const StepArgs &a, int tile_x, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_record(const float *p, const int *rec_cell, float *rec, int rec_ld, int row, int nrec, hipStream_t s);
void launch_c2(const float *vel, float *c2, int ny, int nx, int ldc, float dt, float dx, hipStream_t s);
void launch_stack(float *img, int ny, int nx_w, int x0, double *sum, int ld_sum, int *fold, hipStream_t s);
void launch_normalize(const double *sum, const int *fold, float *out, size_t n, hipStream_t s);
void launch_gather(const float *traces, int nt, const int *cols, int ncols, float *out, hipStream_t s);

}  // namespace psdm
