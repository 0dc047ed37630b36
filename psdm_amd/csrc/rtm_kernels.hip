// gfx950 kernels of the RTM time step: 6th-order acoustic stencil, CPML strip
// recursions, source injection, snapshot store and zero-lag imaging fused in one launch.
//
// Semantics follow the reference solver step by step (solver_ac_forward_ram.f90:344-577,
// solver_ac_backward_ram.f90:329-545): derivatives with clamped edges
// (solver_space_derivs_clean.f90), CPML memory recursions (solver_PML_matrices.f90),
// interior update (:443-444), free surface (:463-473), snapshot (:476-500) and imaging
// (solver_ac_backward_ram.f90:455-481).  Every product and sum is rounded separately in
// the reference's order (built with -ffp-contract=off), so one step reproduces the
// reference FP32 arithmetic.
#include <hip/hip_ext.h>

#include "rtm_core.hpp"

namespace psdm {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// c2 of padded cell (j,i): the reference fills the pads with the nearest edge
// velocity of ctmp (solver_ac_forward_ram.f90:282-309), except that its _ram loops
// fill the top and bottom pads only over the first dn interior columns
// (`do j=1,dnx`, :288-302); the rest of those pads keeps c = 0.
__device__ __forceinline__ float c2_at(const StepArgs &a, int j, int i) {
    const int iw = i - a.dn;
    const bool pad_row = j < a.dn || j >= a.dn + a.ny;
    if (a.ram && pad_row && iw >= a.dn && iw < a.nx) return 0.f;
    const int jw = clampi(j - a.dn, 0, a.ny - 1);
    return a.c2[(size_t)clampi(iw, 0, a.nx - 1) * a.ldc + jw];
}

// Source term at an interior cell (injection :348-358; duplicates resolved on the host
// to the last writer).
__device__ __forceinline__ bool source_at(const StepArgs &a, int ev_s, int ev_e, int cell, float &src) {
    bool has = false;
    for (int e = ev_s; e < ev_e; ++e)
        if (a.ev_cell[e] == cell) {
            src = a.inj_val[(size_t)(a.k - 2) * a.ninj + a.ev_inj[e]];
            has = true;
        }
    return has;
}

__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                           // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace

// One time step of one 64 x TX tile (tile index t) of one shot.  All threads of the
// workgroup take part in the barriers.
template <int TX>
__device__ __forceinline__ void step_tile(const StepArgs &a, int t, float *tile, float *red) {
    constexpr int TY = kTileY, HY = TY + 6, HX = TX + 6, CPW = TX / 4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tj = t % a.ntile_y, ti = t / a.ntile_y;
    const int j0 = tj * TY, i0 = ti * TX;
    const int dn = a.dn;
    const int j = j0 + lane;
    const bool rowok = j < a.nyt;
    const bool interior_tile = j0 >= dn && j0 + TY <= dn + a.ny && i0 >= dn && i0 + TX <= dn + a.nx;

    // loads that do not depend on the staged tile are issued before the barrier
    float p1v[CPW], c2v[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int i = i0 + w * CPW + q;
        const bool ok = rowok && i < a.nxt;
        p1v[q] = (ok && !a.zp) ? a.pn[(size_t)i * a.ld + j] : 0.f;
        c2v[q] = ok ? (interior_tile ? a.c2[(size_t)(i - dn) * a.ldc + (j - dn)] : c2_at(a, j, i)) : 0.f;
    }

    // stage p(k-1) rows j0-3..j0+66, columns i0-3..i0+TX+2 (zero outside the grid)
    for (int li = w; li < HX; li += 4) {
        const int gi = i0 - 3 + li;
        const bool colok = gi >= 0 && gi < a.nxt && !a.zc;
        const float *col = a.pc + (size_t)(colok ? gi : 0) * a.ld;
        int gj = j0 - 3 + lane;
        tile[li * HY + lane] = (colok && gj >= 0 && gj < a.nyt) ? col[gj] : 0.f;
        if (lane < 6) {
            gj = j0 + 61 + lane;
            tile[li * HY + 64 + lane] = (colok && gj < a.nyt) ? col[gj] : 0.f;
        }
    }
    __syncthreads();

    // d = p(k) - p(k-2) of this thread's interior cells, for the snapshot / imaging pass
    float dv[CPW];
    bool dok[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        dv[q] = 0.f;
        dok[q] = false;
    }
    const int ev_s = a.ev_ptr[t], ev_e = a.ev_ptr[t + 1];

#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int i = i0 + w * CPW + q;
        if (!rowok || i >= a.nxt) continue;
        const size_t idx = (size_t)i * a.ld + j;
        const float *T = tile + (i - i0 + 3) * HY + (lane + 3);  // T[di*HY + dj]
        const float p1 = p1v[q], c2 = c2v[q];
        if (interior_tile) {
            const float p = T[0];
            float d2x = (p + p) * kC2_0;
            d2x = d2x + (T[HY] + T[-HY]) * kC2_1;
            d2x = d2x + (T[2 * HY] + T[-2 * HY]) * kC2_2;
            d2x = d2x + (T[3 * HY] + T[-3 * HY]) * kC2_3;
            float d2y = (p + p) * kC2_0;
            d2y = d2y + (T[1] + T[-1]) * kC2_1;
            d2y = d2y + (T[2] + T[-2]) * kC2_2;
            d2y = d2y + (T[3] + T[-3]) * kC2_3;
            float lap = d2x + d2y;
            float src;
            if (ev_s < ev_e && source_at(a, ev_s, ev_e, j + i * a.nyt, src)) lap = lap + src;
            float v = (-p1 + 2.f * p) + lap * c2;  // :443-444
            if (a.fs && j == dn) v = 0.f;          // :463-466
            a.pn[idx] = v;
            dv[q] = v - p1;
            dok[q] = true;
            continue;
        }

        // ---- general cell: CPML strips, corners, clamped derivative edges ----
        const int xs = i < dn ? 1 : (i >= dn + a.nx ? 2 : 0);
        const int ys = j < dn ? 1 : (j >= dn + a.ny ? 2 : 0);
        if (a.fs && ys == 1) {  // with a free surface the top CPML is never updated (:391)
            a.pn[idx] = 0.f;
            continue;
        }
        auto P = [&](int gj, int gi) -> float {
            const int lj = gj - j0 + 3, li = gi - i0 + 3;
            if (lj >= 0 && lj < HY && li >= 0 && li < HX) return tile[li * HY + lj];
            return a.zc ? 0.f : a.pc[(size_t)gi * a.ld + gj];
        };
        // derivatives at the clamped column/row (solver_space_derivs_clean.f90:24-34)
        const int ic = clampi(i, 3, a.nxt - 4), jc = clampi(j, 3, a.nyt - 4);
        const float px = P(j, ic), py = P(jc, i);
        const float xm1 = P(j, ic - 1), xp1 = P(j, ic + 1), xm2 = P(j, ic - 2), xp2 = P(j, ic + 2);
        const float xm3 = P(j, ic - 3), xp3 = P(j, ic + 3);
        const float ym1 = P(jc - 1, i), yp1 = P(jc + 1, i), ym2 = P(jc - 2, i), yp2 = P(jc + 2, i);
        const float ym3 = P(jc - 3, i), yp3 = P(jc + 3, i);
        float d2x = (px + px) * kC2_0;
        d2x = d2x + (xp1 + xm1) * kC2_1;
        d2x = d2x + (xp2 + xm2) * kC2_2;
        d2x = d2x + (xp3 + xm3) * kC2_3;
        float d2y = (py + py) * kC2_0;
        d2y = d2y + (yp1 + ym1) * kC2_1;
        d2y = d2y + (yp2 + ym2) * kC2_2;
        d2y = d2y + (yp3 + ym3) * kC2_3;
        const float p = T[0];
        if (xs == 0 && ys == 0) {
            float lap = d2x + d2y;
            float src;
            if (ev_s < ev_e && source_at(a, ev_s, ev_e, j + i * a.nyt, src)) lap = lap + src;
            float v = (-p1 + 2.f * p) + lap * c2;
            if (a.fs && j == dn) v = 0.f;
            a.pn[idx] = v;
            dv[q] = v - p1;
            dok[q] = true;
            continue;
        }
        const size_t ld = (size_t)a.ld;
        const int dn2 = 2 * dn;
        const int distx = xs == 0 ? 0 : (xs == 1 ? dn - i : i - (dn + a.nx) + 1);  // 1 next to the interior
        const int disty = ys == 0 ? 0 : (ys == 1 ? dn - j : j - (dn + a.ny) + 1);
        // corners use F(max(dist_x, dist_y)) for both directions (solver_PML_coeffs.f90:288-322)
        const int dist = distx > disty ? distx : disty;
        const float A = a.fA[dist - 1], B = a.fB[dist - 1];
        // x-direction memory (strips xl/xr and the corners), pml_matrices_x/_xycorner(1)
        float propx = 0.f;
        if (xs != 0) {
            const int il = xs == 1 ? i : i - (dn + a.nx);  // column inside the strip
            const int xc = xs == 1 ? i : i - a.nx;         // column of the x memory arrays
            const int xcc = xc - il + clampi(il, 3, dn - 4);
            float dphi = 0.f, pho = 0.f, cho = 0.f;
            const size_t xo = (size_t)xc * ld + j;
            if (!a.zm) {
                const float *ph = a.phx_o + j;
                dphi = (ph[(xcc + 1) * ld] - ph[(xcc - 1) * ld]) * kC1_1;
                dphi = dphi + (ph[(xcc + 2) * ld] - ph[(xcc - 2) * ld]) * kC1_2;
                dphi = dphi + (ph[(xcc + 3) * ld] - ph[(xcc - 3) * ld]) * kC1_3;
                pho = a.phx_o[xo];
                cho = a.chx[xo];
            }
            float d1x = (xp1 - xm1) * kC1_1;
            d1x = d1x + (xp2 - xm2) * kC1_2;
            d1x = d1x + (xp3 - xm3) * kC1_3;
            const float phn = A * pho + B * d1x;
            const float chn = A * cho + B * (d2x + dphi);
            a.phx_n[xo] = phn;
            a.chx[xo] = chn;
            propx = dphi + chn;
        }
        float propy = 0.f;
        if (ys != 0) {
            const int jl = ys == 1 ? j : j - (dn + a.ny);
            const int yr = ys == 1 ? j : j - a.ny;  // row of the y memory arrays
            const int yrc = yr - jl + clampi(jl, 3, dn - 4);
            float dphi = 0.f, pho = 0.f, cho = 0.f;
            const size_t yo = (size_t)i * dn2 + yr;
            if (!a.zm) {
                const float *ph = a.phy_o + (size_t)i * dn2;
                dphi = (ph[yrc + 1] - ph[yrc - 1]) * kC1_1;
                dphi = dphi + (ph[yrc + 2] - ph[yrc - 2]) * kC1_2;
                dphi = dphi + (ph[yrc + 3] - ph[yrc - 3]) * kC1_3;
                pho = a.phy_o[yo];
                cho = a.chy[yo];
            }
            float d1y = (yp1 - ym1) * kC1_1;
            d1y = d1y + (yp2 - ym2) * kC1_2;
            d1y = d1y + (yp3 - ym3) * kC1_3;
            const float phn = A * pho + B * d1y;
            const float chn = A * cho + B * (d2y + dphi);
            a.phy_n[yo] = phn;
            a.chy[yo] = chn;
            propy = dphi + chn;
        }
        float v;
        if (ys == 0) {
            // x strip, interior rows (:376-382).  With a free surface the _ram form
            // zeroes row dn+1 of the grid but the strip keeps its own history there.
            const bool own = a.ram && a.fs && j == dn;
            const int hsi = xs == 1 ? i : i - a.nx;
            const float ps = own ? (a.zc ? 0.f : a.hs_c[hsi]) : p;
            const float p1s = own ? (a.zp ? 0.f : a.hs_n[hsi]) : p1;
            v = (-p1s + 2.f * ps) + c2 * ((d2x + d2y) + propx);
            if (own) {
                a.hs_n[hsi] = v;
                v = 0.f;
            }
        } else if (xs == 0) {
            v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propy);  // y strip (:388-396)
        } else {
            v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propx);  // corner x call (:402-403)
            v = v + c2 * propy;                                 // corner y call (:407)
        }
        a.pn[idx] = v;
    }

    if (a.snap_slot < 0) return;  // uniform over the workgroup
    const size_t sbase = (size_t)a.snap_slot * a.snap_stride;
    if (a.mode == 0) {
        // snapshot F = ((p(k) - p(k-2)) * 0.5) / dt (:485)
        float fv[CPW];
#pragma unroll
        for (int q = 0; q < CPW; ++q) fv[q] = dok[q] ? (dv[q] * 0.5f) / a.dt : 0.f;
        if (a.snap_fmt == kSnapF32) {
#pragma unroll
            for (int q = 0; q < CPW; ++q)
                if (dok[q]) {
                    const int i = i0 + w * CPW + q;
                    ((float *)a.snaps)[sbase + (size_t)(i - dn) * a.ny + (j - dn)] = fv[q];
                }
            return;
        }
        // 16-bit snapshot: power-of-two scale from the tile's max-abs
        float m = 0.f;
#pragma unroll
        for (int q = 0; q < CPW; ++q) m = fmaxf(m, fabsf(fv[q]));
        m = wave_max(m);
        if (lane == 0) red[w] = m;
        __syncthreads();
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        __syncthreads();  // red is reused by the next tile
        const int target = a.snap_fmt == kSnapF16 ? 14 : 0;
        const int e = m > 0.f ? ilogbf(m) - target : 0;
        if (threadIdx.x == 0) a.snap_scale[(size_t)a.snap_slot * a.ntiles + t] = ldexpf(1.f, e);
#pragma unroll
        for (int q = 0; q < CPW; ++q)
            if (dok[q]) {
                const int i = i0 + w * CPW + q;
                const size_t so = sbase + (size_t)(i - dn) * a.ny + (j - dn);
                const float s = ldexpf(fv[q], -e);
                if (a.snap_fmt == kSnapF16)
                    ((_Float16 *)a.snaps)[so] = (_Float16)s;
                else
                    ((unsigned short *)a.snaps)[so] = f32_to_bf16(s);
            }
        return;
    }
    // imaging: kernel += 0.5*(q(k)-q(k-2)) * F(snap) * store_snap (solver_ac_backward_ram.f90:465)
    const float scale = a.snap_fmt == kSnapF32 ? 1.f : a.snap_scale[(size_t)a.snap_slot * a.ntiles + t];
#pragma unroll
    for (int q = 0; q < CPW; ++q)
        if (dok[q]) {
            const int i = i0 + w * CPW + q;
            const size_t o = (size_t)(i - dn) * a.ny + (j - dn);
            float F;
            if (a.snap_fmt == kSnapF32)
                F = ((const float *)a.snaps)[sbase + o];
            else if (a.snap_fmt == kSnapF16)
                F = (float)((const _Float16 *)a.snaps)[sbase + o] * scale;
            else
                F = bf16_to_f32(((const unsigned short *)a.snaps)[sbase + o]) * scale;
            const float term = ((0.5f * dv[q]) * F) * a.ss_f;
            if (a.kahan) {
                const float img = a.image[o];
                const float y = term - a.image_comp[o];
                const float tt = img + y;
                a.image_comp[o] = (tt - img) - y;
                a.image[o] = tt;
            } else {
                a.image[o] = a.image[o] + term;
            }
        }
}

template <int TX>
__global__ __launch_bounds__(kBlock) void rtm_step_kernel(StepArgs a) {
    __shared__ float tile[(TX + 6) * (kTileY + 6)];
    __shared__ float red[4];
    const int b = blockIdx.x;
    // traces of the previous step (:567-571), read from p(k-1) before it is overwritten
    if (a.rec_row >= 0)
        for (int r = b * kBlock + threadIdx.x; r < a.nrec; r += gridDim.x * kBlock) {
            const int c = a.rec_cell[r];
            a.rec[(size_t)r * a.rec_ld + a.rec_row] = c >= 0 ? a.pc[c] : 0.f;
        }
    // XCD-aware order: the 8 XCDs take blocks round-robin, so XCD x gets the
    // contiguous tile range [x*per_xcd, (x+1)*per_xcd) and shares halos in its own L2.
    const int t = (b & 7) * a.per_xcd + (b >> 3);
    if (t >= a.ntiles) return;
    step_tile<TX>(a, t, tile, red);
}

// Per-step view of a slot at global step G (forward steps k = 2..nt, then backward
// steps k = 2..nt; ping-pong buffers alternate with k; each phase starts from zero state).
__device__ void resolve_slot(const SlotDesc &sd, long long G, StepArgs *out) {
    const unsigned *src = (const unsigned *)&sd.a;
    unsigned *dst = (unsigned *)out;
    for (int q = 0; q < (int)(sizeof(StepArgs) / 4); ++q) dst[q] = src[q];
    const long long tl = G - sd.g0;
    const int nt = sd.nt;
    if (sd.g0 < 0 || tl < 0 || tl >= 2LL * (nt - 1)) {
        out->k = 0;  // idle
        return;
    }
    const bool fwd = tl < nt - 1;
    const int k = (int)(fwd ? tl + 2 : tl - (nt - 1) + 2);
    const int m = (k - 2) & 1;
    out->k = k;
    out->mode = fwd ? 0 : 1;
    out->pc = m ? sd.pB : sd.pA;
    out->pn = m ? sd.pA : sd.pB;
    out->phx_o = m ? sd.phxB : sd.phxA;
    out->phx_n = m ? sd.phxA : sd.phxB;
    out->phy_o = m ? sd.phyB : sd.phyA;
    out->phy_n = m ? sd.phyA : sd.phyB;
    out->hs_c = m ? sd.hsB : sd.hsA;
    out->hs_n = m ? sd.hsA : sd.hsB;
    out->zc = k == 2;
    out->zp = k <= 3;
    out->zm = k == 2;
    out->rec_row = -1;
    int slot = -1;
    if (fwd) {
        if ((k - 1) % sd.ss == 0) {
            const int s = (k - 1) / sd.ss;
            if (s <= sd.nsnap) slot = s - 1;
        }
        out->ev_ptr = sd.ev_ptr_f;
        out->ev_cell = sd.ev_cell_f;
        out->ev_inj = sd.ev_inj_f;
        out->inj_val = sd.inj_f;
        out->ninj = sd.ninj_f;
    } else {
        // backward step k pairs with the forward snapshot taken at k_f = nt + 2 - k
        if ((nt + 1 - k) % sd.ss == 0) {
            const int s = (nt + 1 - k) / sd.ss;
            if (s >= 1 && s <= sd.nsnap) slot = s - 1;
        }
        out->ev_ptr = sd.ev_ptr_b;
        out->ev_cell = sd.ev_cell_b;
        out->ev_inj = sd.ev_inj_b;
        out->inj_val = sd.inj_b;
        out->ninj = sd.ninj_b;
    }
    out->snap_slot = slot;
}

// One time step of every active slot: a persistent grid walks the slot x tile work
// list; each XCD takes a contiguous range of items (neighbouring tiles of one shot).
// Block 0 also resolves the next step's per-slot view into cur[(G+1)&1].
template <int TX>
__global__ __launch_bounds__(kBlock) void rtm_batch_kernel(BatchArgs b, StepArgs *cur) {
    __shared__ float tile[(TX + 6) * (kTileY + 6)];
    __shared__ float red[4];
    const long long G = *b.gbase + b.goff;
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < b.nslots; s += blockDim.x)
            resolve_slot(b.slots[s], G + 1, &cur[((G + 1) & 1) * b.nslots + s]);
    const StepArgs *now = cur + (G & 1) * b.nslots;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    for (int m = blockIdx.x >> 3; m < b.per_xcd; m += nbx) {
        const int wk = x * b.per_xcd + m;
        if (wk >= b.nwork) break;
        const int s = wk / b.cap_tiles, t = wk - s * b.cap_tiles;
        const StepArgs &a = now[s];
        if (a.k == 0 || t >= a.ntiles) continue;
        step_tile<TX>(a, t, tile, red);
        __syncthreads();  // the tile buffer is reused by the next item
    }
}

// Resolves the per-slot views of step G (used before the first step of a chunk).
__global__ void rtm_resolve_kernel(const SlotDesc *slots, int nslots, const long long *gbase, StepArgs *cur) {
    const long long G = *gbase;
    for (int s = threadIdx.x; s < nslots; s += blockDim.x) resolve_slot(slots[s], G, &cur[(G & 1) * nslots + s]);
}

__global__ void rtm_advance_kernel(long long *gbase, long long by) {
    if (threadIdx.x == 0) *gbase += by;
}

__global__ void rtm_record_kernel(const float *p, const int *rec_cell, float *rec, int rec_ld, int row, int nrec) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) {
        const int c = rec_cell[r];
        rec[(size_t)r * rec_ld + row] = c >= 0 ? p[c] : 0.f;
    }
}

// c2 = c**2*dt**2/dmodel**2 (solver_ac_forward_ram.f90:319), FP32 in the same order.
__global__ void rtm_c2_kernel(const float *vel, float *c2, int ny, int nx, int ldc, float dt, float dx) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j < ny && i < nx) {
        const float c = vel[(size_t)i * ny + j];
        c2[(size_t)i * ldc + j] = ((c * c) * (dt * dt)) / (dx * dx);
    }
}

// out[r*ncols + e] = traces[cols[e]*nt + r'] with r' = r or nt-1-r (time reversal of
// the gather, BackSource = Data(nt:1:-1), calculate_PSDM_RTM.f90:448-450)
__global__ void rtm_gather_kernel(const float *traces, int nt, const int *cols, int ncols, float *out, int reverse) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (e < ncols && r < nt - 1) out[(size_t)r * ncols + e] = traces[(size_t)cols[e] * nt + (reverse ? nt - 1 - r : r)];
}

// Cross-shot stack (calculate_PSDM_RTM.f90:471-497): the shot image kernel(ny,nx_w)
// is embedded at columns x0..x0+nx_w-1 of PSDM_sum (FP64) and the fold Cov_sum gains 1
// there (Cov_w = 1 on the window).  The shot image (and its compensation term) is
// cleared for the next shot (the driver's PSDM_w = 0, :384).
__global__ void rtm_stack_kernel(float *img, float *comp, int ny, int nx_w, int x0, double *sum, int ld_sum,
                                 int *fold) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j < ny && i < nx_w) {
        const size_t o = (size_t)i * ny + j;
        const size_t g = (size_t)(x0 + i) * ld_sum + j;
        double v = (double)img[o];
        if (comp) {
            v -= (double)comp[o];  // Kahan: the running sum is img - comp
            comp[o] = 0.f;
        }
        sum[g] += v;
        fold[g] += 1;
        img[o] = 0.f;
    }
}

// normalize_psdm (auxiliar_PSDM.f90:1-20): PSDM = PSDM_sum / Cov_sum where the fold is
// positive, 0 elsewhere.
__global__ void rtm_normalize_kernel(const double *sum, const int *fold, float *out, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = fold[e] > 0 ? (float)(sum[e] / (double)fold[e]) : 0.f;
}

// hipExtLaunchKernelGGL stamps start/stop on the dispatch packet itself, so a sampled
// step reports the kernel's own duration (as a profiler's kernel trace does).
void launch_step(const StepArgs &a, int tile_x, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const int grid = 8 * a.per_xcd;
    if (tile_x == 16)
        hipExtLaunchKernelGGL(rtm_step_kernel<16>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, a);
    else
        hipExtLaunchKernelGGL(rtm_step_kernel<8>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, a);
}

void launch_batch(const BatchArgs &b, StepArgs *cur, int tile_x, int grid, hipStream_t s, hipEvent_t e0,
                  hipEvent_t e1) {
    grid = (grid + 7) / 8 * 8;
    if (!e0 && !e1) {  // plain launch: also valid under stream capture
        if (tile_x == 16)
            hipLaunchKernelGGL(rtm_batch_kernel<16>, dim3(grid), dim3(kBlock), 0, s, b, cur);
        else
            hipLaunchKernelGGL(rtm_batch_kernel<8>, dim3(grid), dim3(kBlock), 0, s, b, cur);
        return;
    }
    if (tile_x == 16)
        hipExtLaunchKernelGGL(rtm_batch_kernel<16>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, b, cur);
    else
        hipExtLaunchKernelGGL(rtm_batch_kernel<8>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, b, cur);
}

void launch_resolve(const SlotDesc *slots, int nslots, const long long *gbase, StepArgs *cur, hipStream_t s) {
    hipLaunchKernelGGL(rtm_resolve_kernel, dim3(1), dim3(256), 0, s, slots, nslots, gbase, cur);
}

void launch_advance(long long *gbase, long long by, hipStream_t s) {
    hipLaunchKernelGGL(rtm_advance_kernel, dim3(1), dim3(64), 0, s, gbase, by);
}

void launch_record(const float *p, const int *rec_cell, float *rec, int rec_ld, int row, int nrec, hipStream_t s) {
    if (nrec <= 0) return;
    hipLaunchKernelGGL(rtm_record_kernel, dim3((nrec + 255) / 256), dim3(256), 0, s, p, rec_cell, rec, rec_ld, row,
                       nrec);
}

void launch_c2(const float *vel, float *c2, int ny, int nx, int ldc, float dt, float dx, hipStream_t s) {
    hipLaunchKernelGGL(rtm_c2_kernel, dim3((ny + 255) / 256, nx), dim3(256), 0, s, vel, c2, ny, nx, ldc, dt, dx);
}

void launch_gather(const float *traces, int nt, const int *cols, int ncols, float *out, int reverse, hipStream_t s) {
    if (ncols <= 0 || nt < 2) return;
    hipLaunchKernelGGL(rtm_gather_kernel, dim3((ncols + 63) / 64, nt - 1), dim3(64), 0, s, traces, nt, cols, ncols,
                       out, reverse);
}

void launch_stack(float *img, float *comp, int ny, int nx_w, int x0, double *sum, int ld_sum, int *fold,
                  hipStream_t s) {
    hipLaunchKernelGGL(rtm_stack_kernel, dim3((ny + 255) / 256, nx_w), dim3(256), 0, s, img, comp, ny, nx_w, x0, sum,
                       ld_sum, fold);
}

void launch_normalize(const double *sum, const int *fold, float *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(rtm_normalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sum, fold, out, n);
}

}  // namespace psdm
