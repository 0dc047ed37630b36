// Two time steps per launch for the batched engine (temporal blocking).
//
// One launch takes a shot from levels (k-1, k-2) of the wavefield and the CPML memories to
// levels (k+1, k): the same per-step arithmetic as the one-step kernels (the reference's
// solver_ac_forward_ram.f90:344-577 / solver_ac_backward_ram.f90:329-545 step, every
// product and sum rounded separately in its order), evaluated twice inside one work item.
// No item depends on another within a launch: each computes the first step redundantly
// over a halo wide enough for the second (the work map is described with TbTiling in
// rtm_core.hpp), reads only the launch's input buffers and writes only its output buffers.
// Per cell, the wavefield crosses HBM once per two steps instead of once per step.
//
// FP32 snapshots and imaging only (plain or Kahan): the 16-bit formats keep one scale per
// 248-row stream tile, which the two-step work map does not have.
#include <hip/hip_ext.h>

#include "rtm_core.hpp"
#include "rtm_device.hpp"

namespace psdm {

namespace {

// Per-launch fields of one slot (resolved on the device from the global launch index).
struct TbStep {
    int k0;      // first time step of the launch (reference index k)
    int two;     // 1: steps k0 and k0+1; 0: k0 only (the last launch of a phase of odd length)
    int mode;    // 0 forward (snapshots), 1 backward (imaging)
    int first;   // first launch of the phase: every input level and memory reads as zero
    const float *pl, *pp;  // in: p(k0-1), p(k0-2)
    float *ol, *op;        // out: p(k0+1) and p(k0) (with two = 0: p(k0) in ol)
    const float *phx_i, *chx_i, *phy_i, *chy_i;
    float *phx_o, *chx_o, *phy_o, *chy_o;
    const float *hs_l, *hs_p;  // x-strip history of the free-surface row, in: levels k0-1, k0-2
    float *hs_ol, *hs_op;      // out: as ol / op
    int snap[2];               // snapshot slot of steps k0 and k0+1 (-1 none)
    const int *ev_ptr, *ev_cell, *ev_inj;
    const float *inj_val;
    int ninj;
};

// Launch l of a slot (l = G - g0): launches 0 .. M-1 run the forward steps 2..nt two at a
// time (M = nt/2, the last one single when nt-1 is odd), launches M .. 2M-1 the backward
// steps.  Level pairs alternate: even launches of a phase read (pA, pB) and write (pC,
// pD), odd ones the reverse; the CPML memories alternate between their A and B buffers.
__device__ __forceinline__ bool resolve_tb(const SlotDesc &sd, long long G, TbStep &s) {
    const long long tl = G - sd.g0;
    const int nt = sd.nt, M = nt / 2;
    if (sd.g0 < 0 || tl < 0 || tl >= 2LL * M) return false;
    const bool fwd = tl < M;
    const int m = (int)(fwd ? tl : tl - M);
    const bool even = (m & 1) == 0;
    s.k0 = 2 + 2 * m;
    s.two = s.k0 + 1 <= nt;
    s.mode = fwd ? 0 : 1;
    s.first = m == 0;
    s.pl = even ? sd.pA : sd.pC;
    s.pp = even ? sd.pB : sd.pD;
    s.ol = even ? sd.pC : sd.pA;
    s.op = even ? sd.pD : sd.pB;
    s.phx_i = even ? sd.phxA : sd.phxB;
    s.phx_o = even ? sd.phxB : sd.phxA;
    s.chx_i = even ? sd.a.chx : sd.chxB;
    s.chx_o = even ? sd.chxB : sd.a.chx;
    s.phy_i = even ? sd.phyA : sd.phyB;
    s.phy_o = even ? sd.phyB : sd.phyA;
    s.chy_i = even ? sd.a.chy : sd.chyB;
    s.chy_o = even ? sd.chyB : sd.a.chy;
    s.hs_l = even ? sd.hsA : sd.hsC;
    s.hs_p = even ? sd.hsB : sd.hsD;
    s.hs_ol = even ? sd.hsC : sd.hsA;
    s.hs_op = even ? sd.hsD : sd.hsB;
    for (int q = 0; q < 2; q++) {
        const int k = s.k0 + q;
        int slot = -1;
        if (fwd) {
            const int sn = fdiv(k - 1, sd.fd_ss);
            if (sn * sd.ss == k - 1 && sn <= sd.nsnap) slot = sn - 1;
        } else {  // backward step k pairs with the forward snapshot of step nt + 2 - k
            const int sn = fdiv(nt + 1 - k, sd.fd_ss);
            if (sn * sd.ss == nt + 1 - k && sn >= 1 && sn <= sd.nsnap) slot = sn - 1;
        }
        s.snap[q] = (q == 1 && !s.two) ? -1 : slot;
    }
    s.ev_ptr = fwd ? sd.tev_ptr_f : sd.tev_ptr_b;
    s.ev_cell = fwd ? sd.tev_cell_f : sd.tev_cell_b;
    s.ev_inj = fwd ? sd.tev_inj_f : sd.tev_inj_b;
    s.inj_val = fwd ? sd.inj_f : sd.inj_b;
    s.ninj = fwd ? sd.ninj_f : sd.ninj_b;
    return true;
}

// ------------------------------------------------------------------------------------
// Edge tiles (CPML strips and the interior band along them), in LDS.

constexpr int kLH = 72;                 // LDS rows: owned rows (<= 52) + 2 * kTbH
constexpr int kLW = kTbTX + 2 * kTbH;   // LDS columns: 34
constexpr int kCR = 64, kCC = kTbTX + 12;  // cells a tile computes: rows +-6 (one per lane), columns +-6
constexpr int kEB = 512, kEW = kEB / 64;   // edge-tile workgroup: 8 wavefronts
constexpr int kCPW = (kCC + kEW - 1) / kEW;  // columns per wavefront (4)
constexpr int kTabMax = 256;            // CPML cells per side the edge kernel supports (dPML + 6)
// LDS of an edge tile (floats): p, phi_x, phi_y of the current level; the injection values of
// the two steps and the event flags per computed cell; the damping tables
constexpr int kEdgeLds = 3 * kLH * kLW + 3 * kCR * kCC + 2 * kTabMax;

// Level n+1 of padded cell (j, i) from level n: p, phi_x, phi_y of level n at grid positions
// in LDS (column-major, kLH rows from row lj0, columns from li0); p1 (level n-1), c2 and
// the cell's chi memories (in/out) from the caller.  The formulas are those of the
// one-step kernels (cpml_tile and stream_item in rtm_kernels.hip), cell type by cell type.
// wantP: the new p (and chi) are needed; otherwise only the new phi (a cell of the ring
// that only the next step's memory derivative reads).  tA, tB: the damping tables in LDS.
struct EdgeView {
    const float *P, *X, *Y;
    int lj0, li0;
    __device__ __forceinline__ float p(int j, int i) const { return P[(i - li0) * kLH + (j - lj0)]; }
    __device__ __forceinline__ float x(int j, int i) const { return X[(i - li0) * kLH + (j - lj0)]; }
    __device__ __forceinline__ float y(int j, int i) const { return Y[(i - li0) * kLH + (j - lj0)]; }
};

__device__ __forceinline__ void edge_cell(const StepArgs &a, const EdgeView &L, const float *tA, const float *tB, int j,
                                          int i, bool wantP, float p1, float c2, bool has_src, float src, float &chx,
                                          float &chy, float &hsl, float &hsp, float &np, float &nphx, float &nphy) {
    const int dn = a.dn;
    const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0), ys = j < dn ? 1 : (j >= dn + a.ny ? 2 : 0);
    const float p = L.p(j, i);
    if (xs == 0 && ys == 0) {  // interior (stream_item)
        if (!wantP) return;
        float lap = d2(p, L.p(j, i + 1), L.p(j, i - 1), L.p(j, i + 2), L.p(j, i - 2), L.p(j, i + 3), L.p(j, i - 3)) +
                    d2(p, L.p(j + 1, i), L.p(j - 1, i), L.p(j + 2, i), L.p(j - 2, i), L.p(j + 3, i), L.p(j - 3, i));
        if (has_src) lap = lap + src;
        float v = (-p1 + 2.f * p) + lap * c2;
        if (a.fs && j == dn) v = 0.f;
        np = v;
        return;
    }
    // the top CPML with a free surface, and the _ram top/bottom pads past column dn (c2 = 0):
    // p stays 0 and their memories are never updated (nor read)
    if ((a.fs && ys == 1) || (a.ram && ys != 0 && xs == 0 && i - dn >= dn)) {
        np = 0.f;
        return;
    }
    const int jc = clampi(j, 3, a.nyt - 4), ic = clampi(i, 3, a.nxt - 4);
    const int distx = xs == 0 ? 0 : (xs == 1 ? dn - i : i - (dn + a.nx) + 1);
    const int disty = ys == 0 ? 0 : (ys == 1 ? dn - j : j - (dn + a.ny) + 1);
    const int dist = distx > disty ? distx : disty;  // corners: F(max(dist_x, dist_y))
    const float A = tA[dist - 1], B = tB[dist - 1];
    const float px = L.p(j, ic), xp1 = L.p(j, ic + 1), xm1 = L.p(j, ic - 1), xp2 = L.p(j, ic + 2),
                xm2 = L.p(j, ic - 2), xp3 = L.p(j, ic + 3), xm3 = L.p(j, ic - 3);
    const float py = L.p(jc, i), yp1 = L.p(jc + 1, i), ym1 = L.p(jc - 1, i), yp2 = L.p(jc + 2, i),
                ym2 = L.p(jc - 2, i), yp3 = L.p(jc + 3, i), ym3 = L.p(jc - 3, i);
    float d2x = (px + px) * kC2_0;
    d2x = d2x + (xp1 + xm1) * kC2_1;
    d2x = d2x + (xp2 + xm2) * kC2_2;
    d2x = d2x + (xp3 + xm3) * kC2_3;
    float d2y = (py + py) * kC2_0;
    d2y = d2y + (yp1 + ym1) * kC2_1;
    d2y = d2y + (yp2 + ym2) * kC2_2;
    d2y = d2y + (yp3 + ym3) * kC2_3;
    float propx = 0.f, propy = 0.f;
    if (xs != 0) {  // pml_matrices_x / _xycorner (x call)
        const int xoff = xs == 2 ? a.nx : 0, xlo = xs == 2 ? dn : 0;
        float d1x = (xp1 - xm1) * kC1_1;
        d1x = d1x + (xp2 - xm2) * kC1_2;
        d1x = d1x + (xp3 - xm3) * kC1_3;
        nphx = A * L.x(j, i) + B * d1x;
        if (wantP) {  // (a ring cell's memory derivative would reach past the staged halo)
            const int icc = xoff + xlo + clampi(i - xoff - xlo, 3, dn - 4);  // clamped derivative column (grid)
            float dphi = (L.x(j, icc + 1) - L.x(j, icc - 1)) * kC1_1;
            dphi = dphi + (L.x(j, icc + 2) - L.x(j, icc - 2)) * kC1_2;
            dphi = dphi + (L.x(j, icc + 3) - L.x(j, icc - 3)) * kC1_3;
            chx = A * chx + B * (d2x + dphi);
            propx = dphi + chx;
        }
    }
    if (ys != 0) {  // pml_matrices_y / _xycorner (y call)
        const int ylo = ys == 2 ? dn + a.ny : 0;
        float d1y = (yp1 - ym1) * kC1_1;
        d1y = d1y + (yp2 - ym2) * kC1_2;
        d1y = d1y + (yp3 - ym3) * kC1_3;
        nphy = A * L.y(j, i) + B * d1y;
        if (wantP) {
            const int jcc = ylo + clampi(j - ylo, 3, dn - 4);
            float dphi = (L.y(jcc + 1, i) - L.y(jcc - 1, i)) * kC1_1;
            dphi = dphi + (L.y(jcc + 2, i) - L.y(jcc - 2, i)) * kC1_2;
            dphi = dphi + (L.y(jcc + 3, i) - L.y(jcc - 3, i)) * kC1_3;
            chy = A * chy + B * (d2y + dphi);
            propy = dphi + chy;
        }
    }
    if (!wantP) return;
    float v;
    if (ys == 0) {
        // x strip, interior rows (:376-382); with a free surface the _ram form zeroes row dn+1
        // of the grid but the strip keeps its own history there
        const bool own = a.ram && a.fs && j == dn;
        const float ps = own ? hsl : p, p1s = own ? hsp : p1;
        v = (-p1s + 2.f * ps) + c2 * ((d2x + d2y) + propx);
        if (own) {
            hsp = hsl;
            hsl = v;
            v = 0.f;
        }
    } else if (xs == 0) {
        v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propy);  // y strip (:388-396)
    } else {
        v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propx);  // corner x call (:402-403)
        v = v + c2 * propy;                                 // corner y call (:407)
    }
    np = v;
}

template <int SV>
__device__ void tb_edge_tile(const SlotDesc &sd, const TbStep &s, int e, float *lds) {
    constexpr bool kKahan = SV == kSvF32Kahan;
    const StepArgs &a = sd.a;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dn = a.dn, nyt = a.nyt, nxt = a.nxt;
    const size_t ld = a.ld, dn2 = 2 * (size_t)dn;
    int j0, j1, i0, i1;
    tb_edge_tile(sd.tb, e, a.ny, a.nx, dn, &j0, &j1, &i0, &i1);
    float *sP = lds, *sX = lds + kLH * kLW, *sY = lds + 2 * kLH * kLW;
    float *sV0 = lds + 3 * kLH * kLW, *sV1 = sV0 + kCR * kCC;
    int *evm = (int *)(sV1 + kCR * kCC);
    float *tA = (float *)(evm + kCR * kCC), *tB = tA + kTabMax;
    const int lj0 = j0 - kTbH, li0 = i0 - kTbH;
    const bool zin = s.first != 0;

    // stage level k0-1 of p and of the CPML memories over the tile +-9 (zero outside the
    // grid), the damping tables, and clear the event flags
    for (int c = w; c < kLW; c += kEW) {
        const int i = li0 + c;
        const bool colok = i >= 0 && i < nxt;
        const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0);
        const int xc = i - (xs == 2 ? a.nx : 0);
        for (int r = lane; r < kLH; r += 64) {
            const int j = lj0 + r;
            const bool ok = colok && j >= 0 && j < nyt && !zin;
            const int ys = j < dn ? 1 : (j >= dn + a.ny ? 2 : 0);
            sP[c * kLH + r] = ok ? gp(s.pl)[(size_t)i * ld + j] : 0.f;
            sX[c * kLH + r] = ok && xs != 0 ? gp(s.phx_i)[(size_t)xc * ld + j] : 0.f;
            sY[c * kLH + r] = ok && ys != 0 ? gp(s.phy_i)[(size_t)i * dn2 + (ys == 1 ? j : j - a.ny)] : 0.f;
        }
    }
    for (int q = threadIdx.x; q < dn; q += kEB) {
        tA[q] = gp(a.fA)[q];
        tB[q] = gp(a.fB)[q];
    }
    for (int q = threadIdx.x; q < kCR * kCC; q += kEB) evm[q] = -1;
    __syncthreads();
    // injection cells of this tile (its step-k0 cells +-3): both steps' values by computed cell
    const int ev0 = gp(s.ev_ptr)[e], ev1 = gp(s.ev_ptr)[e + 1];
    gptr<const float> inj0 = gp(s.inj_val) + (size_t)(s.k0 - 2) * s.ninj;
    gptr<const float> inj1 = gp(s.inj_val) + (size_t)(s.k0 - 1) * s.ninj;
    for (int q = ev0 + (int)threadIdx.x; q < ev1; q += kEB) {
        const int cell = gp(s.ev_cell)[q], ij = gp(s.ev_inj)[q];
        const int m = (dn + (cell >> 16) - (i0 - 6)) * kCR + (dn + (cell & 0xffff) - (j0 - 6));
        evm[m] = q;
        sV0[m] = inj0[ij];
        sV1[m] = s.two ? inj1[ij] : 0.f;
    }
    __syncthreads();

    // this thread's cells: row j0-6+lane, columns i0-6+w+kEW*q (q < kCPW)
    const int j = j0 - 6 + lane;
    const bool rowin = j >= 0 && j < nyt && j < j1 + 6;
    const int ys = j < dn ? 1 : (j >= dn + a.ny ? 2 : 0);
    float p1[kCPW], c2v[kCPW], chx[kCPW], chy[kCPW], hsl[kCPW], hsp[kCPW], np1[kCPW], nx1[kCPW], ny1[kCPW], tm1[kCPW];
    const EdgeView L{sP, sX, sY, lj0, li0};
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
        const int i = i0 - 6 + w + kEW * q;
        const bool in = rowin && w + kEW * q < kCC && i >= 0 && i < nxt && i < i1 + 6;
        const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0);
        const bool near = j >= j0 - 3 && j < j1 + 3 && i >= i0 - 3 && i < i1 + 3;  // p computed at step k0
        p1[q] = (in && near && !zin) ? gp(s.pp)[(size_t)i * ld + j] : 0.f;
        c2v[q] = (in && near) ? c2_at(a, j, i) : 0.f;
        const int xc = i - (xs == 2 ? a.nx : 0);
        chx[q] = (in && near && xs != 0 && !zin) ? gp(s.chx_i)[(size_t)xc * ld + j] : 0.f;
        chy[q] = (in && near && ys != 0 && !zin) ? gp(s.chy_i)[(size_t)i * dn2 + (ys == 1 ? j : j - a.ny)] : 0.f;
        const bool own = in && a.ram && a.fs && j == dn && xs != 0;
        hsl[q] = (own && !zin) ? gp(s.hs_l)[xc] : 0.f;
        hsp[q] = (own && !zin) ? gp(s.hs_p)[xc] : 0.f;
        np1[q] = nx1[q] = ny1[q] = tm1[q] = 0.f;
    }
    const bool imaging = s.mode == 1;

    // step k0: p and chi over the tile +-3, phi over the tile +-6
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
        const int i = i0 - 6 + w + kEW * q;
        if (!(rowin && w + kEW * q < kCC && i >= 0 && i < nxt && i < i1 + 6)) continue;
        const bool near = j >= j0 - 3 && j < j1 + 3 && i >= i0 - 3 && i < i1 + 3;
        const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0);
        const int m = (w + kEW * q) * kCR + lane;
        nx1[q] = L.x(j, i);
        ny1[q] = L.y(j, i);
        edge_cell(a, L, tA, tB, j, i, near, p1[q], c2v[q], evm[m] >= 0, sV0[m], chx[q], chy[q], hsl[q], hsp[q],
                  np1[q], nx1[q], ny1[q]);
        const bool owned = j >= j0 && j < j1 && i >= i0 && i < i1;
        if (owned && xs == 0 && ys == 0 && s.snap[0] >= 0) {
            const size_t co = (size_t)(i - dn) * a.lds + (j - dn);
            const float dv = np1[q] - p1[q];
            if (!imaging) {
                __builtin_nontemporal_store((dv * 0.5f) / a.dt, gp((float *)a.snaps) + (size_t)s.snap[0] * a.snap_stride + co);
            } else {
                const float F = __builtin_nontemporal_load(gp((const float *)a.snaps) + (size_t)s.snap[0] * a.snap_stride + co);
                tm1[q] = ((0.5f * dv) * F) * a.ss_f;
            }
        }
    }
    if (s.two) {
        __syncthreads();  // every cell of level k0-1 has been read
#pragma unroll
        for (int q = 0; q < kCPW; ++q) {
            const int i = i0 - 6 + w + kEW * q;
            if (!(rowin && w + kEW * q < kCC && i >= 0 && i < nxt && i < i1 + 6)) continue;
            const int c = i - li0, r = j - lj0;
            const bool near = j >= j0 - 3 && j < j1 + 3 && i >= i0 - 3 && i < i1 + 3;
            const float pold = sP[c * kLH + r];
            if (near) sP[c * kLH + r] = np1[q];
            sX[c * kLH + r] = nx1[q];
            sY[c * kLH + r] = ny1[q];
            p1[q] = pold;  // level k0-1: the previous level of step k0+1
        }
        __syncthreads();
    }
    // step k0+1 on the owned cells (or, with a single step, step k0's results), and the
    // outputs: the last two levels of p, the memories and the history, the image
#pragma unroll
    for (int q = 0; q < kCPW; ++q) {
        const int i = i0 - 6 + w + kEW * q;
        if (!(j >= j0 && j < j1 && i >= i0 && i < i1)) continue;
        const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0);
        const int m = (w + kEW * q) * kCR + lane;
        float np2 = np1[q], nx2 = nx1[q], ny2 = ny1[q], tm2 = 0.f;
        if (s.two) {
            nx2 = L.x(j, i);
            ny2 = L.y(j, i);
            edge_cell(a, L, tA, tB, j, i, true, p1[q], c2v[q], evm[m] >= 0, sV1[m], chx[q], chy[q], hsl[q], hsp[q],
                      np2, nx2, ny2);
            if (xs == 0 && ys == 0 && s.snap[1] >= 0) {
                const size_t co = (size_t)(i - dn) * a.lds + (j - dn);
                const float dv = np2 - p1[q];
                if (!imaging) {
                    __builtin_nontemporal_store((dv * 0.5f) / a.dt,
                                                gp((float *)a.snaps) + (size_t)s.snap[1] * a.snap_stride + co);
                } else {
                    const float F =
                        __builtin_nontemporal_load(gp((const float *)a.snaps) + (size_t)s.snap[1] * a.snap_stride + co);
                    tm2 = ((0.5f * dv) * F) * a.ss_f;
                }
            }
        }
        const size_t o = (size_t)i * ld + j;
        const int xc = i - (xs == 2 ? a.nx : 0);
        gp(s.ol)[o] = np2;
        if (s.two) gp(s.op)[o] = np1[q];
        if (xs != 0) {
            gp(s.phx_o)[(size_t)xc * ld + j] = nx2;
            gp(s.chx_o)[(size_t)xc * ld + j] = chx[q];
        }
        if (ys != 0) {
            const size_t yo = (size_t)i * dn2 + (ys == 1 ? j : j - a.ny);
            gp(s.phy_o)[yo] = ny2;
            gp(s.chy_o)[yo] = chy[q];
        }
        if (a.ram && a.fs && j == dn && xs != 0) {
            gp(s.hs_ol)[xc] = hsl[q];
            gp(s.hs_op)[xc] = hsp[q];
        }
        if (xs == 0 && ys == 0 && imaging && (s.snap[0] >= 0 || s.snap[1] >= 0)) {
            // the image gets step k0's term, then step k0+1's (stream_item's sum or Kahan update)
            const size_t co = (size_t)(i - dn) * a.lds + (j - dn);
            float im = gp(a.image)[co], cm = kKahan ? gp(a.image_comp)[co] : 0.f;
            const float terms[2] = {tm1[q], tm2};
            const int has[2] = {s.snap[0] >= 0, s.snap[1] >= 0};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!has[h]) continue;
                if (kKahan) {
                    const float y = terms[h] - cm, t = im + y;
                    cm = (t - im) - y;
                    im = t;
                } else {
                    im = im + terms[h];
                }
            }
            gp(a.image)[co] = im;
            if (kKahan) gp(a.image_comp)[co] = cm;
        }
    }
}

// ------------------------------------------------------------------------------------
// Core items: wavefront stream items over the interior away from the CPML.

// Injection events of one column for one step: adds the values of the item's events at
// column `col` (ordered by column, then row) to the lanes holding their rows; `e` is the
// step's cursor.  The first 64 events of the item are held one per lane (ev_c, ev_v0 / ev_v1:
// the two steps' values) and read with readlane.
__device__ __forceinline__ void tb_inject(v4f &lap, int col, int &e, int ev_e, int e_first, int e_reg, int ev_c,
                                          float ev_v, gptr<const int> ev_cell, gptr<const int> ev_inj,
                                          gptr<const float> inj, int row0, int lane) {
    while (e < ev_e) {
        int cell;
        float val;
        if (e < e_reg) {
            cell = __builtin_amdgcn_readlane(ev_c, e - e_first);
            if ((cell >> 16) != col) break;
            val = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ev_v), e - e_first));
        } else {
            cell = ev_cell[e];
            if ((cell >> 16) != col) break;
            val = inj[ev_inj[e]];
        }
        const int r = (cell & 0xffff) - row0;
        if (lane == (r >> 2)) {
            const int q = r & 3;
            if (q == 0) lap.x = lap.x + val;
            else if (q == 1) lap.y = lap.y + val;
            else if (q == 2) lap.z = lap.z + val;
            else lap.w = lap.w + val;
        }
        ++e;
    }
}

// Laplacian of the lane's 4 rows at the centre column w3 (x neighbours w0..w6, y neighbours
// from the lanes on either side), in the reference's order: d2x + d2y.
__device__ __forceinline__ v4f lap4(const v4f &w0, const v4f &w1, const v4f &w2, const v4f &w3, const v4f &w4,
                                    const v4f &w5, const v4f &w6) {
    const float b1 = from_prev_lane(w3.y), b2 = from_prev_lane(w3.z), b3 = from_prev_lane(w3.w);
    const float n0 = from_next_lane(w3.x), n1 = from_next_lane(w3.y), n2 = from_next_lane(w3.z);
    v4f lap;
    lap.x = d2(w3.x, w4.x, w2.x, w5.x, w1.x, w6.x, w0.x) + d2(w3.x, w3.y, b3, w3.z, b2, w3.w, b1);
    lap.y = d2(w3.y, w4.y, w2.y, w5.y, w1.y, w6.y, w0.y) + d2(w3.y, w3.z, w3.x, w3.w, b3, n0, b2);
    lap.z = d2(w3.z, w4.z, w2.z, w5.z, w1.z, w6.z, w0.z) + d2(w3.z, w3.w, w3.y, n0, w3.x, n1, b3);
    lap.w = d2(w3.w, w4.w, w2.w, w5.w, w1.w, w6.w, w0.w) + d2(w3.w, n0, w3.z, n1, w3.y, n2, w3.x);
    return lap;
}

__device__ __forceinline__ v4f upd4(const v4f &p1, const v4f &p, const v4f &lap, const v4f &cc) {
    v4f v;
    v.x = (-p1.x + 2.f * p.x) + lap.x * cc.x;  // :443-444
    v.y = (-p1.y + 2.f * p.y) + lap.y * cc.y;
    v.z = (-p1.z + 2.f * p.z) + lap.z * cc.z;
    v.w = (-p1.w + 2.f * p.w) + lap.w * cc.w;
    return v;
}

// Image accumulation of one step's term (stream_item's plain sum or Kahan update).
template <bool kKahan>
__device__ __forceinline__ void img_add(v4f &im, v4f &cm, const v4f &t) {
    if (kKahan) {
        v4f y, s;
        y.x = t.x - cm.x;
        y.y = t.y - cm.y;
        y.z = t.z - cm.z;
        y.w = t.w - cm.w;
        s.x = im.x + y.x;
        s.y = im.y + y.y;
        s.z = im.z + y.z;
        s.w = im.w + y.w;
        cm.x = (s.x - im.x) - y.x;
        cm.y = (s.y - im.y) - y.y;
        cm.z = (s.z - im.z) - y.z;
        cm.w = (s.w - im.w) - y.w;
        im = s;
    } else {
        im.x = im.x + t.x;
        im.y = im.y + t.y;
        im.z = im.z + t.z;
        im.w = im.w + t.w;
    }
}

// Stores v at the owned rows of a lane (all 4, some or none).
__device__ __forceinline__ void store_owned(float *base, unsigned lo, const v4f &v, bool full, bool any, bool ow0,
                                            bool ow1, bool ow2, bool ow3) {
    if (full) {
        *lanep<v4f>(base, lo) = v;
    } else if (any) {
        gptr<float> q = lanep<float>(base, lo);
        if (ow0) q[0] = v.x;
        if (ow1) q[1] = v.y;
        if (ow2) q[2] = v.z;
        if (ow3) q[3] = v.w;
    }
}

// Core item u: interior rows o-8 .. o+247 (4 per lane; lanes 2..61 own the rows of their
// kTbRows-row tile) and the columns [c0, cend).  Step k0 runs 3 columns ahead of step
// k0+1: at column t it forms p(k0) from the window of p(k0-1) columns t-3..t+3, then
// p(k0+1) at column t-3 from the window of p(k0) columns t-6..t.  Step k0 covers the
// columns c0-3 .. cend+2, step k0+1 the owned columns; both steps' rows with their halo
// stay inside the interior, so every value they read is one this item computed or an
// input level.  Backward: a column's image is read once, gets the two steps' terms in
// order, and is written once.
template <int SV>
__device__ void tb_core_item(const SlotDesc &sd, const TbStep &s, int u) {
    constexpr bool kKahan = SV == kSvF32Kahan;
    const StepArgs &a = sd.a;
    const int lane = threadIdx.x & 63;
    const int dn = a.dn, cr = sd.tb.cr;
    const int CS = u / cr, R = u - CS * cr;
    const int o = tb_origin(R, a.ny);
    const int own0 = kTbTop + kTbRows * R, own1 = own0 + kTbRows < a.ny - 3 ? own0 + kTbRows : a.ny - 3;
    const int c0 = 3 + CS * sd.tb.sw, cend = c0 + sd.tb.sw < a.nx - 3 ? c0 + sd.tb.sw : a.nx - 3;
    const int jl = o - 8 + 4 * lane;  // interior row of this lane's first value
    const bool ow0 = jl >= own0 && jl < own1, ow1 = jl + 1 >= own0 && jl + 1 < own1;
    const bool ow2 = jl + 2 >= own0 && jl + 2 < own1, ow3 = jl + 3 >= own0 && jl + 3 < own1;
    const bool full = ow0 && ow3, any = ow0 || ow1 || ow2 || ow3;
    const unsigned lo = 16u * (unsigned)lane;
    const size_t ld = a.ld;
    const int prow = dn + o - 8;  // padded row of lane 0's first value
    const float *P = s.pl + prow, *Q = s.pp + prow;
    float *N1 = (s.two ? s.op : s.ol) + prow, *N2 = s.ol + prow;
    const float *C2 = a.c2 + (o - 8);
    const v4f z4 = v4f{0.f, 0.f, 0.f, 0.f};
    const bool zin = s.first != 0;
    const bool imaging = s.mode == 1, two = s.two != 0;
    const float *SN0 = (const float *)a.snaps + (size_t)(s.snap[0] < 0 ? 0 : s.snap[0]) * a.snap_stride + (o - 8);
    const float *SN1 = (const float *)a.snaps + (size_t)(s.snap[1] < 0 ? 0 : s.snap[1]) * a.snap_stride + (o - 8);
    float *IM = a.image + (o - 8), *CM = a.image_comp + (o - 8);
    // interior row 0 (free surface) in this item's rows: lane fl, component fq
    const int fr = -(o - 8), fl = a.fs && fr >= 0 && fr < 256 ? fr >> 2 : -1, fq = fr & 3;

    // injection events of the item (its step-k0 columns c0-3 .. cend+2, all its rows),
    // ordered by column then row; the first 64 held one per lane
    const int eb = sd.tb.ne + u;
    int e1 = gp(s.ev_ptr)[eb];
    const int ev_e = gp(s.ev_ptr)[eb + 1];
    int e2 = e1;
    const int e_first = e1, e_reg = ev_e < e1 + 64 ? ev_e : e1 + 64;
    gptr<const float> inj0 = gp(s.inj_val) + (size_t)(s.k0 - 2) * s.ninj;
    gptr<const float> inj1 = gp(s.inj_val) + (size_t)(s.k0 - 1) * s.ninj;
    int ev_c = 0;
    float ev_v0 = 0.f, ev_v1 = 0.f;
    if (e1 + lane < e_reg) {
        ev_c = gp(s.ev_cell)[e1 + lane];
        const int ij = gp(s.ev_inj)[e1 + lane];
        ev_v0 = inj0[ij];
        ev_v1 = two ? inj1[ij] : 0.f;
    }
    // step k0+1 starts at column c0: its cursor skips the events of the halo columns
    while (e2 < ev_e && ((e2 < e_reg ? __builtin_amdgcn_readlane(ev_c, e2 - e_first) : gp(s.ev_cell)[e2]) >> 16) < c0) ++e2;

    // window of p(k0-1): columns t-3 .. t+3 (W1[0..6]); of p(k0): columns t-6 .. t (W2[0..6])
    v4f a0, a1, a2, a3, a4, a5, a6;
    v4f b0 = z4, b1 = z4, b2 = z4, b3 = z4, b4 = z4, b5 = z4, b6 = z4;
    if (zin) {
        a0 = a1 = a2 = a3 = a4 = a5 = z4;
    } else {
        a0 = ldv(P + (size_t)(dn + c0 - 6) * ld, lo);
        a1 = ldv(P + (size_t)(dn + c0 - 5) * ld, lo);
        a2 = ldv(P + (size_t)(dn + c0 - 4) * ld, lo);
        a3 = ldv(P + (size_t)(dn + c0 - 3) * ld, lo);
        a4 = ldv(P + (size_t)(dn + c0 - 2) * ld, lo);
        a5 = ldv(P + (size_t)(dn + c0 - 1) * ld, lo);
    }
    // terms of step k0 for the 3 columns before the one step k0+1 images next
    v4f tq0 = z4, tq1 = z4, tq2 = z4;
    for (int t = c0 - 3; t < cend + 3; ++t) {
        a6 = zin ? z4 : ldv(P + (size_t)(dn + t + 3) * ld, lo);
        const v4f p1 = zin ? z4 : ldv(Q + (size_t)(dn + t) * ld, lo);
        const v4f cc = ldv(C2 + (size_t)t * a.ldc, lo);
        // step k0 at column t
        v4f lap = lap4(a0, a1, a2, a3, a4, a5, a6);
        tb_inject(lap, t, e1, ev_e, e_first, e_reg, ev_c, ev_v0, gp(s.ev_cell), gp(s.ev_inj), inj0, o - 8, lane);
        v4f v = upd4(p1, a3, lap, cc);
        if (lane == fl) {
            if (fq == 0) v.x = 0.f;
            else if (fq == 1) v.y = 0.f;
            else if (fq == 2) v.z = 0.f;
            else v.w = 0.f;
        }
        const bool own_t = t >= c0 && t < cend;
        v4f tn = z4;  // step k0's imaging term of column t
        if (own_t) {
            store_owned(N1 + (size_t)(dn + t) * ld, lo, v, full, any, ow0, ow1, ow2, ow3);
            const size_t co = (size_t)t * a.lds;
            v4f dv;
            dv.x = v.x - p1.x;
            dv.y = v.y - p1.y;
            dv.z = v.z - p1.z;
            dv.w = v.w - p1.w;
            if (s.snap[0] >= 0) {
                if (!imaging) {
                    v4f f;
                    f.x = (dv.x * 0.5f) / a.dt;
                    f.y = (dv.y * 0.5f) / a.dt;
                    f.z = (dv.z * 0.5f) / a.dt;
                    f.w = (dv.w * 0.5f) / a.dt;
                    if (full) {
                        __builtin_nontemporal_store(f, lanep<v4f>(SN0 + co, lo));
                    } else if (any) {
                        gptr<float> q = lanep<float>(SN0 + co, lo);
                        if (ow0) q[0] = f.x;
                        if (ow1) q[1] = f.y;
                        if (ow2) q[2] = f.z;
                        if (ow3) q[3] = f.w;
                    }
                } else {
                    const v4f F = __builtin_nontemporal_load(lanep<const v4f>(SN0 + co, lo));
                    tn.x = ((0.5f * dv.x) * F.x) * a.ss_f;
                    tn.y = ((0.5f * dv.y) * F.y) * a.ss_f;
                    tn.z = ((0.5f * dv.z) * F.z) * a.ss_f;
                    tn.w = ((0.5f * dv.w) * F.w) * a.ss_f;
                }
            }
            if (imaging && !two && any && s.snap[0] >= 0) {  // a single step: its term goes straight in
                v4f im = ldv(IM + co, lo), cm = kKahan ? ldv(CM + co, lo) : z4;
                img_add<kKahan>(im, cm, tn);
                store_owned(IM + co, lo, im, full, any, ow0, ow1, ow2, ow3);
                if (kKahan) store_owned(CM + co, lo, cm, full, any, ow0, ow1, ow2, ow3);
            }
        }
        // window of p(k0): columns t-6 .. t
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = b4;
        b4 = b5;
        b5 = b6;
        b6 = v;
        // step k0+1 at column t-3
        const int t2 = t - 3;
        if (two && t2 >= c0) {
            const v4f cc2 = ldv(C2 + (size_t)t2 * a.ldc, lo);
            v4f lap2 = lap4(b0, b1, b2, b3, b4, b5, b6);
            tb_inject(lap2, t2, e2, ev_e, e_first, e_reg, ev_c, ev_v1, gp(s.ev_cell), gp(s.ev_inj), inj1, o - 8, lane);
            v4f v2 = upd4(a0, b3, lap2, cc2);  // p(k0-1) at column t-3 is a0
            if (lane == fl) {
                if (fq == 0) v2.x = 0.f;
                else if (fq == 1) v2.y = 0.f;
                else if (fq == 2) v2.z = 0.f;
                else v2.w = 0.f;
            }
            store_owned(N2 + (size_t)(dn + t2) * ld, lo, v2, full, any, ow0, ow1, ow2, ow3);
            const size_t co = (size_t)t2 * a.lds;
            v4f dv;
            dv.x = v2.x - a0.x;
            dv.y = v2.y - a0.y;
            dv.z = v2.z - a0.z;
            dv.w = v2.w - a0.w;
            if (!imaging) {
                if (s.snap[1] >= 0) {
                    v4f f;
                    f.x = (dv.x * 0.5f) / a.dt;
                    f.y = (dv.y * 0.5f) / a.dt;
                    f.z = (dv.z * 0.5f) / a.dt;
                    f.w = (dv.w * 0.5f) / a.dt;
                    if (full) {
                        __builtin_nontemporal_store(f, lanep<v4f>(SN1 + co, lo));
                    } else if (any) {
                        gptr<float> q = lanep<float>(SN1 + co, lo);
                        if (ow0) q[0] = f.x;
                        if (ow1) q[1] = f.y;
                        if (ow2) q[2] = f.z;
                        if (ow3) q[3] = f.w;
                    }
                }
            } else if (any && (s.snap[0] >= 0 || s.snap[1] >= 0)) {
                // the column's image: step k0's term (computed 3 columns ago), then step k0+1's
                v4f im = ldv(IM + co, lo), cm = kKahan ? ldv(CM + co, lo) : z4;
                if (s.snap[0] >= 0) img_add<kKahan>(im, cm, tq0);
                if (s.snap[1] >= 0) {
                    const v4f F = __builtin_nontemporal_load(lanep<const v4f>(SN1 + co, lo));
                    v4f t1;
                    t1.x = ((0.5f * dv.x) * F.x) * a.ss_f;
                    t1.y = ((0.5f * dv.y) * F.y) * a.ss_f;
                    t1.z = ((0.5f * dv.z) * F.z) * a.ss_f;
                    t1.w = ((0.5f * dv.w) * F.w) * a.ss_f;
                    img_add<kKahan>(im, cm, t1);
                }
                store_owned(IM + co, lo, im, full, any, ow0, ow1, ow2, ow3);
                if (kKahan) store_owned(CM + co, lo, cm, full, any, ow0, ow1, ow2, ow3);
            }
        }
        tq0 = tq1;
        tq1 = tq2;
        tq2 = tn;
        a0 = a1;
        a1 = a2;
        a2 = a3;
        a3 = a4;
        a4 = a5;
        a5 = a6;
    }
}

// Slot and item of work item wk of a launch group's list: every slot holding `uni` items
// (uni > 0), or else the compact list of the occupied slots (offsets woff per slot), as in
// rtm_batch_kernel (rtm_kernels.hip).
template <bool CMP, int LIST>
__device__ __forceinline__ void tb_locate(const SlotDesc *slots, int n, int wk, int uni, const FastDiv &wfd,
                                          const int *woffs, int *s, int *t) {
    if (!CMP || uni > 0) {
        *s = fdiv(wk, wfd);
        *t = wk - *s * uni;
    } else if (n <= 64) {
        const int mine = woffs[threadIdx.x & 63];
        *s = __popcll(__ballot(mine <= wk)) - 1;
        *t = wk - __builtin_amdgcn_readlane(mine, *s);
    } else {
        int q = min(fdiv(wk, wfd), n - 1);
        auto off = [&](int k) { return LIST == 0 ? slots[k].woff : slots[k].cwoff; };
        while (q > 0 && off(q) > wk) --q;
        while (q + 1 < n && off(q + 1) <= wk) ++q;
        *s = q;
        *t = wk - off(q);
    }
}

// One launch (two time steps) of the edge tiles of every active slot of a launch group
// (list 0: SlotDesc woff / wtotal / wuni / wfd); each XCD takes a contiguous range.
template <int SV, bool CMP>
__global__ __launch_bounds__(kEB) void rtm_tb_edge_kernel(BatchArgs b, const SlotDesc *__restrict__ slots,
                                                          const long long *__restrict__ gbase) {
    __shared__ __attribute__((aligned(16))) float lds[kEdgeLds];
    __shared__ int woffs[64];
    const long long G = *gbase + b.goff;
    const int n = b.nslots;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    int total = n * b.cap_tiles, uni = b.cap_tiles;
    FastDiv wfd = b.fd_cap;
    if (CMP) {
        total = slots[0].wtotal;
        uni = slots[0].wuni;
        wfd = slots[0].wfd;
        if (n <= 64 && (int)threadIdx.x < 64) woffs[threadIdx.x] = (int)threadIdx.x < n ? slots[threadIdx.x].woff : 0x7fffffff;
        __syncthreads();
    }
    const int per_xcd = (total + 7) >> 3;
    for (int m = blockIdx.x >> 3; m < per_xcd; m += nbx) {
        const int wk = x * per_xcd + m;
        if (wk >= total) break;
        int s, t;
        tb_locate<CMP, 0>(slots, n, wk, uni, wfd, woffs, &s, &t);
        const SlotDesc &sd = slots[s];
        TbStep st;
        if (!resolve_tb(sd, G, st) || t >= sd.tb.ne) continue;
        tb_edge_tile<SV>(sd, st, t, lds);
        __syncthreads();  // the LDS is reused by the next tile
    }
}

// One launch (two time steps) of the core items of every active slot of a launch group,
// four per workgroup (list 1: SlotDesc cwoff / cwtotal / cwuni / cwfd).
template <int SV, bool CMP>
__global__ __launch_bounds__(kBlock) void rtm_tb_core_kernel(BatchArgs b, const SlotDesc *__restrict__ slots,
                                                             const long long *__restrict__ gbase) {
    __shared__ int woffs[64];
    const long long G = *gbase + b.goff;
    const int n = b.nslots;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    int total = n * b.cap_tiles, uni = b.cap_tiles;
    FastDiv wfd = b.fd_cap;
    if (CMP) {
        total = slots[0].cwtotal;
        uni = slots[0].cwuni;
        wfd = slots[0].cwfd;
        if (n <= 64 && (int)threadIdx.x < 64) woffs[threadIdx.x] = (int)threadIdx.x < n ? slots[threadIdx.x].cwoff : 0x7fffffff;
        __syncthreads();
    }
    const int per_xcd = (total + 7) >> 3;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int m = blockIdx.x >> 3; m < per_xcd; m += nbx) {
        const int wk = x * per_xcd + m;
        if (wk >= total) break;
        int s, t;
        tb_locate<CMP, 1>(slots, n, wk, uni, wfd, woffs, &s, &t);
        const SlotDesc &sd = slots[s];
        TbStep st;
        if (!resolve_tb(sd, G, st) || t >= sd.tb.ncblk) continue;
        const int u = 4 * t + w;
        if (u < sd.tb.ncore) tb_core_item<SV>(sd, st, u);
    }
}

template <int SV>
void launch_tb_sv(const BatchArgs &b, int part, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    auto k = part == 0 ? (b.compact ? rtm_tb_edge_kernel<SV, true> : rtm_tb_edge_kernel<SV, false>)
                       : (b.compact ? rtm_tb_core_kernel<SV, true> : rtm_tb_core_kernel<SV, false>);
    const int block = part == 0 ? kEB : kBlock;
    if (!e0 && !e1)
        hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, s, b, b.slots, b.gbase);
    else
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(block), 0, s, e0, e1, 0, b, b.slots, b.gbase);
}

}  // namespace

void launch_tb_batch(const BatchArgs &b, int part, int sv, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    grid = (grid + 7) / 8 * 8;
    if (sv == kSvF32Kahan)
        launch_tb_sv<kSvF32Kahan>(b, part, grid, s, e0, e1);
    else
        launch_tb_sv<kSvF32>(b, part, grid, s, e0, e1);
}

}  // namespace psdm
