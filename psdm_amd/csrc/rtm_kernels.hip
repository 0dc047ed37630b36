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
#include "rtm_device.hpp"

namespace psdm {

// Stages p(k-1) rows j0-3..j0+66 and columns i0-3..i0+TX+2 of the padded grid into
// LDS tile[(i-i0+3)*(TY+6) + (j-j0+3)] (zero outside the grid or when `zero`).
template <int TX>
__device__ __forceinline__ void stage_p(gptr<const float> pc, int ld, int nyt, int nxt, int j0, int i0, bool zero,
                                        float *tile) {
    constexpr int HY = kTileY + 6, HX = TX + 6;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // w: wave-uniform
#pragma unroll
    for (int li = w; li < HX; li += 4) {
        const int gi = i0 - 3 + li;
        const bool colok = gi >= 0 && gi < nxt && !zero;
        gptr<const float> col = pc + (colok ? gi : 0) * ld;
        int gj = j0 - 3 + lane;
        tile[li * HY + lane] = (colok && gj >= 0 && gj < nyt) ? col[gj] : 0.f;
        if (lane < 6) {
            gj = j0 + 61 + lane;
            tile[li * HY + 64 + lane] = (colok && gj < nyt) ? col[gj] : 0.f;
        }
    }
}

// Interior stream item u: interior rows o-4 .. o+251 of one stream row tile (4 per lane;
// lanes 0 and 63 hold the 4-row halos) and columns [c0, cend), walked left to right with
// p(k-1) columns i-3 .. i+3 in registers.  x neighbours come from the register window,
// y neighbours from the lane's own 4 rows and its two neighbour lanes (DPP).  Per column:
// update (:443-444, source injection :348-358), free surface (:463-466), then the
// snapshot (forward, :476-500) or the imaging term (backward,
// solver_ac_backward_ram.f90:455-481).  Only the rows the item owns are stored.
template <int SV>
__device__ __forceinline__ void stream_item(const StepArgs &a, int u, v4f *stage) {
    constexpr int SB = SV & 3;                                // the SnapVariant proper
    constexpr bool kF32 = SB == kSvF32 || SB == kSvF32Kahan;  // else 16-bit (FP16 or BF16)
    constexpr bool kKahan = SB == kSvF32Kahan || SB == kSv16Kahan;
    constexpr bool kZS = (SV & kSvZeroSum) != 0;
    const int lane = threadIdx.x & 63;
    const int dn = a.dn;
    const int CS = fdiv(u, a.fd_ri), R = u - CS * a.ri;
    const int o = stream_origin(R, a.ny);
    const int own0 = R * kStreamRows, own1 = own0 + kStreamRows < a.ny ? own0 + kStreamRows : a.ny;
    const int c0 = dn + CS * a.sw, cend = c0 + a.sw < dn + a.nx ? c0 + a.sw : dn + a.nx;
    const int jl = o - 4 + 4 * lane;  // interior row of this lane's first value
    const bool ow0 = jl >= own0 && jl < own1, ow1 = jl + 1 >= own0 && jl + 1 < own1;
    const bool ow2 = jl + 2 >= own0 && jl + 2 < own1, ow3 = jl + 3 >= own0 && jl + 3 < own1;
    const bool full = ow0 && ow3, any = ow0 || ow1 || ow2 || ow3;
    // Every access is a wave-uniform column base (scalar registers) plus this lane's
    // 32-bit byte offset, so addresses cost one vector register for all arrays.
    const unsigned lo = 16u * (unsigned)lane, lo2 = 8u * (unsigned)lane;
    const int prow = dn + o - 4;  // padded row of lane 0's first value
    const float *P = a.pc + prow;
    float *PN = a.pn + prow;
    const float *C2 = a.c2 + (o - 4);
    const size_t ld = a.ld;
    const v4f z4 = v4f{0.f, 0.f, 0.f, 0.f};

    // injection events of this item, ordered by column then row: cell = iw << 16 | jw; the
    // one-step lists keep event e's values in column e of inj_val.  The item's range comes
    // through the scalar cache; the first 64 events are loaded one per lane (cell and
    // value, independent loads) together with the window, so the column loop reads them
    // with readlane instead of dependent loads.
    int e = sload(a.ev_ptr + u);
    const int ev_e = sload(a.ev_ptr + u + 1);
    gptr<const float> inj = gp(a.inj_val) + (size_t)(a.k - 2) * a.ninj;
    const int e_first = e, e_reg = ev_e < e + 64 ? ev_e : e + 64;
    int ev_c = 0;
    float ev_v = 0.f;
    if (e + lane < e_reg) {
        ev_c = gp(a.ev_cell)[e + lane];
        ev_v = inj[e + lane];
    }

    const bool imaging = a.mode == 1 && a.snap_slot >= 0;
    const bool snap = a.mode == 0 && a.snap_slot >= 0;
    const size_t sbase = (size_t)(a.snap_slot < 0 ? 0 : a.snap_slot) * a.snap_stride;
    const size_t scbase = (size_t)(a.snap_slot < 0 ? 0 : a.snap_slot) * a.scale_stride + (size_t)R * a.nx;
    const float *SN = (const float *)a.snaps + sbase + (o - 4);  // FP32 snapshot, lane 0's first row
    const unsigned short *SH = (const unsigned short *)a.snaps + sbase + (o - 4);
    float *IM = a.image + (o - 4), *CM = a.image_comp + (o - 4);

    // FP32 imaging with at most kStageCols columns: the item's snapshot columns are copied
    // straight into this wavefront's LDS slice (global_load_lds, no registers held) at the
    // start, so their HBM latency is paid once per item instead of once per column.  The
    // compiler does not order these copies against LDS reads; the column loop waits for
    // them (vmcnt 0) before its first read.
    const bool staged = kF32 && imaging && cend - c0 <= kStageCols;
    if (staged)
        for (int q = 0; q < cend - c0; ++q)
            __builtin_amdgcn_global_load_lds((const void *)lanep<const v4f>(SN + (size_t)(c0 - dn + q) * a.lds, lo),
                                             (__attribute__((address_space(3))) void *)(stage + q * 64), 16, 0, 0);

    v4f w0, w1, w2, w3, w4, w5, w6;
    if (a.zc) {
        w0 = w1 = w2 = w3 = w4 = w5 = w6 = z4;
    } else {
        w0 = ldv(P + (c0 - 3) * ld, lo);
        w1 = ldv(P + (c0 - 2) * ld, lo);
        w2 = ldv(P + (c0 - 1) * ld, lo);
        w3 = ldv(P + c0 * ld, lo);
        w4 = ldv(P + (c0 + 1) * ld, lo);
        w5 = ldv(P + (c0 + 2) * ld, lo);
        w6 = ldv(P + (c0 + 3) * ld, lo);
    }
    for (int i = c0; i < cend; ++i) {
        const int iw = i - dn;
        const v4f p1 = a.zp ? z4 : ldv(PN + i * ld, lo);
        const v4f cc = ldv(C2 + (size_t)iw * a.ldc, lo);
        const v4f nw = a.zc ? z4 : ldv(P + (i + 4) * ld, lo);
        const size_t co = (size_t)iw * a.lds;  // column offset in the interior arrays
        v4f F4 = z4, im4 = z4, cm4 = z4;
        float scale = 1.f;
        if (imaging) {
            if (staged) {
                if (i == c0) __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
                F4 = stage[(i - c0) * 64 + lane];
            } else if (kF32) {
                F4 = __builtin_nontemporal_load(lanep<const v4f>(SN + co, lo));
            } else {
                const v2u h = __builtin_nontemporal_load(lanep<const v2u>(SH + co, lo2));
                if (a.snap_fmt == kSnapF16) {
                    F4.x = (float)__builtin_bit_cast(_Float16, (unsigned short)(h.x & 0xffffu));
                    F4.y = (float)__builtin_bit_cast(_Float16, (unsigned short)(h.x >> 16));
                    F4.z = (float)__builtin_bit_cast(_Float16, (unsigned short)(h.y & 0xffffu));
                    F4.w = (float)__builtin_bit_cast(_Float16, (unsigned short)(h.y >> 16));
                } else {
                    F4.x = bf16_to_f32((unsigned short)(h.x & 0xffffu));
                    F4.y = bf16_to_f32((unsigned short)(h.x >> 16));
                    F4.z = bf16_to_f32((unsigned short)(h.y & 0xffffu));
                    F4.w = bf16_to_f32((unsigned short)(h.y >> 16));
                }
                scale = gp(a.snap_scale)[scbase + iw];
            }
            im4 = ldv(IM + co, lo);
            if (kKahan) cm4 = ldv(CM + co, lo);
        }
        // y neighbours: the previous lane's rows 1..3 and the next lane's rows 0..2
        const float b1 = from_prev_lane(w3.y), b2 = from_prev_lane(w3.z), b3 = from_prev_lane(w3.w);
        const float n0 = from_next_lane(w3.x), n1 = from_next_lane(w3.y), n2 = from_next_lane(w3.z);
        v4f lap;
        if (kZS) {
            lap.x = d2z(w3.x, w4.x, w2.x, w5.x, w1.x, w6.x, w0.x) + d2z(w3.x, w3.y, b3, w3.z, b2, w3.w, b1);
            lap.y = d2z(w3.y, w4.y, w2.y, w5.y, w1.y, w6.y, w0.y) + d2z(w3.y, w3.z, w3.x, w3.w, b3, n0, b2);
            lap.z = d2z(w3.z, w4.z, w2.z, w5.z, w1.z, w6.z, w0.z) + d2z(w3.z, w3.w, w3.y, n0, w3.x, n1, b3);
            lap.w = d2z(w3.w, w4.w, w2.w, w5.w, w1.w, w6.w, w0.w) + d2z(w3.w, n0, w3.z, n1, w3.y, n2, w3.x);
        } else {
            lap.x = d2(w3.x, w4.x, w2.x, w5.x, w1.x, w6.x, w0.x) + d2(w3.x, w3.y, b3, w3.z, b2, w3.w, b1);
            lap.y = d2(w3.y, w4.y, w2.y, w5.y, w1.y, w6.y, w0.y) + d2(w3.y, w3.z, w3.x, w3.w, b3, n0, b2);
            lap.z = d2(w3.z, w4.z, w2.z, w5.z, w1.z, w6.z, w0.z) + d2(w3.z, w3.w, w3.y, n0, w3.x, n1, b3);
            lap.w = d2(w3.w, w4.w, w2.w, w5.w, w1.w, w6.w, w0.w) + d2(w3.w, n0, w3.z, n1, w3.y, n2, w3.x);
        }
        // (d2x + d2y + source) at the injection cells of this column (:443)
        while (e < ev_e) {
            int cell;
            float val;
            if (e < e_reg) {
                cell = __builtin_amdgcn_readlane(ev_c, e - e_first);
                if ((cell >> 16) != iw) break;
                val = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ev_v), e - e_first));
            } else {  // events past the first 64 of the item
                cell = gp(a.ev_cell)[e];
                if ((cell >> 16) != iw) break;
                val = inj[e];
            }
            const int r = (cell & 0xffff) - (o - 4);
            if (lane == (r >> 2)) {
                const int q = r & 3;
                if (q == 0) lap.x = lap.x + val;
                else if (q == 1) lap.y = lap.y + val;
                else if (q == 2) lap.z = lap.z + val;
                else lap.w = lap.w + val;
            }
            ++e;
        }
        v4f v;
        v.x = (-p1.x + 2.f * w3.x) + lap.x * cc.x;  // :443-444
        v.y = (-p1.y + 2.f * w3.y) + lap.y * cc.y;
        v.z = (-p1.z + 2.f * w3.z) + lap.z * cc.z;
        v.w = (-p1.w + 2.f * w3.w) + lap.w * cc.w;
        if (a.fs && o == 0) {  // interior row 0 is lane 1's first value (:463-466)
            if (lane == 1) v.x = 0.f;
        }
        if (full) {
            *lanep<v4f>(PN + i * ld, lo) = v;
        } else if (any) {
            gptr<float> q = lanep<float>(PN + i * ld, lo);
            if (ow0) q[0] = v.x;
            if (ow1) q[1] = v.y;
            if (ow2) q[2] = v.z;
            if (ow3) q[3] = v.w;
        }
        v4f dv;
        dv.x = v.x - p1.x;
        dv.y = v.y - p1.y;
        dv.z = v.z - p1.z;
        dv.w = v.w - p1.w;
        if (snap) {
            // F = ((p(k) - p(k-2)) * 0.5) / dt (:485); snapshots are written once and read
            // once, in the backward phase, so they bypass the caches (non-temporal)
            v4f f;
            f.x = (dv.x * 0.5f) / a.dt;
            f.y = (dv.y * 0.5f) / a.dt;
            f.z = (dv.z * 0.5f) / a.dt;
            f.w = (dv.w * 0.5f) / a.dt;
            if (kF32) {
                if (full) {
                    __builtin_nontemporal_store(f, lanep<v4f>(SN + co, lo));
                } else if (any) {
                    gptr<float> q = lanep<float>(SN + co, lo);
                    if (ow0) q[0] = f.x;
                    if (ow1) q[1] = f.y;
                    if (ow2) q[2] = f.z;
                    if (ow3) q[3] = f.w;
                }
            } else {
                // 16-bit: power-of-two scale from the column's max-abs over the owned rows
                float m = 0.f;
                if (ow0) m = fmaxf(m, fabsf(f.x));
                if (ow1) m = fmaxf(m, fabsf(f.y));
                if (ow2) m = fmaxf(m, fabsf(f.z));
                if (ow3) m = fmaxf(m, fabsf(f.w));
                m = wave_max(m);
                const int target = a.snap_fmt == kSnapF16 ? 14 : 0;
                const int ex = m > 0.f ? ilogbf(m) - target : 0;
                if (lane == 0) gp(a.snap_scale)[scbase + iw] = ldexpf(1.f, ex);
                unsigned short hs[4];
                const float fs4[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float sv = ldexpf(fs4[c], -ex);
                    hs[c] = a.snap_fmt == kSnapF16 ? __builtin_bit_cast(unsigned short, (_Float16)sv) : f32_to_bf16(sv);
                }
                if (full) {
                    __builtin_nontemporal_store(v2u{(unsigned)hs[0] | ((unsigned)hs[1] << 16), (unsigned)hs[2] | ((unsigned)hs[3] << 16)}, lanep<v2u>(SH + co, lo2));
                } else if (any) {
                    gptr<unsigned short> q = lanep<unsigned short>(SH + co, lo2);
                    if (ow0) q[0] = hs[0];
                    if (ow1) q[1] = hs[1];
                    if (ow2) q[2] = hs[2];
                    if (ow3) q[3] = hs[3];
                }
            }
        } else if (imaging) {
            // kernel += 0.5*(q(k)-q(k-2)) * F(snap) * store_snap (solver_ac_backward_ram.f90:465)
            v4f F = F4;
            if (!kF32) {
                F.x = F.x * scale;
                F.y = F.y * scale;
                F.z = F.z * scale;
                F.w = F.w * scale;
            }
            v4f term, img, cmp;
            term.x = ((0.5f * dv.x) * F.x) * a.ss_f;
            term.y = ((0.5f * dv.y) * F.y) * a.ss_f;
            term.z = ((0.5f * dv.z) * F.z) * a.ss_f;
            term.w = ((0.5f * dv.w) * F.w) * a.ss_f;
            if (kKahan) {
                v4f y;
                y.x = term.x - cm4.x;
                y.y = term.y - cm4.y;
                y.z = term.z - cm4.z;
                y.w = term.w - cm4.w;
                img.x = im4.x + y.x;
                img.y = im4.y + y.y;
                img.z = im4.z + y.z;
                img.w = im4.w + y.w;
                cmp.x = (img.x - im4.x) - y.x;
                cmp.y = (img.y - im4.y) - y.y;
                cmp.z = (img.z - im4.z) - y.z;
                cmp.w = (img.w - im4.w) - y.w;
            } else {
                img.x = im4.x + term.x;
                img.y = im4.y + term.y;
                img.z = im4.z + term.z;
                img.w = im4.w + term.w;
            }
            if (full) {
                *lanep<v4f>(IM + co, lo) = img;
                if (kKahan) *lanep<v4f>(CM + co, lo) = cmp;
            } else if (any) {
                gptr<float> qi = lanep<float>(IM + co, lo);
                if (ow0) qi[0] = img.x;
                if (ow1) qi[1] = img.y;
                if (ow2) qi[2] = img.z;
                if (ow3) qi[3] = img.w;
                if (kKahan) {
                    gptr<float> qc = lanep<float>(CM + co, lo);
                    if (ow0) qc[0] = cmp.x;
                    if (ow1) qc[1] = cmp.y;
                    if (ow2) qc[2] = cmp.z;
                    if (ow3) qc[3] = cmp.w;
                }
            }
        }
        w0 = w1;
        w1 = w2;
        w2 = w3;
        w3 = w4;
        w4 = w5;
        w5 = w6;
        w6 = nw;
    }
}

// CPML tile: x strip (xs != 0, ys == 0), y strip (xs == 0, ys != 0) or corner.  xs / ys:
// 1 low side, 2 high side.  Every tile spans >= 4 cells along each CPML axis, so all
// clamped-edge stencil points (solver_space_derivs_clean.f90:24-34) and x-memory
// neighbours lie inside the LDS-staged ranges.
template <int TX, bool ZS>
__device__ __forceinline__ void cpml_tile(const StepArgs &a, int xs, int ys, int j0, int j1, int i0, int i1,
                                          float *tile, float *phs) {
    constexpr int TY = kTileY, HY = TY + 6, HX = TX + 6, CPW = TX / 4;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // w: wave-uniform
    const int dn = a.dn, ld = a.ld, dn2 = 2 * dn;
    const int j = j0 + lane;
    const bool rowok = j < j1;
    gptr<float> pn = gp(a.pn);

    // the top CPML is never updated with a free surface (:391); the _ram top and bottom
    // pads past column dn have c2 = 0 (velocity padding :288-302), so p stays exactly 0
    // there and their CPML memory never reaches p.  make_tiling leaves both kinds out of
    // the work list (the solver zeroes both wavefield buffers before a shot); a caller with
    // its own work list that does list such a tile gets its zeros stored in the first two
    // steps of a phase (zp: k <= 3), one into each ping-pong buffer.
    const bool zero_tile = (a.fs && ys == 1) || (a.ram && ys != 0 && xs == 0 && i0 - dn >= dn);
    if (zero_tile) {
        if (rowok && a.zp)
#pragma unroll
            for (int q = 0; q < CPW; ++q) {
                const int i = i0 + w * CPW + q;
                if (i < i1) pn[i * ld + j] = 0.f;
            }
        return;
    }

    // x memory of this strip: columns xm0-3 .. xm0+TX+2 of the memory arrays (strip
    // columns [xlo, xlo+dn)), rows j0..j0+63, staged next to the wavefield tile
    const int xoff = xs == 2 ? a.nx : 0;  // grid column i -> memory column i - xoff
    const int xlo = xs == 2 ? dn : 0;
    const int xm0 = i0 - xoff;
    if (xs != 0) {
        gptr<const float> ph = gp(a.phx_o);
#pragma unroll
        for (int li = w; li < HX; li += 4) {
            const int mc = xm0 - 3 + li;
            const bool ok = rowok && !a.zm && mc >= xlo && mc < xlo + dn;
            phs[li * TY + lane] = ok ? ph[mc * ld + j] : 0.f;
        }
    }
    // every global operand of the update is loaded before the barrier, so a tile costs
    // one memory round trip: p(k-2), c2, the damping factors, the strips' chi (x) and
    // the y memory terms, the _ram strip history
    gptr<const float> fA = gp(a.fA), fB = gp(a.fB);
    const int jc = clampi(j, 3, a.nyt - 4);
    const int disty = ys == 0 ? 0 : (ys == 1 ? dn - j : j - (dn + a.ny) + 1);  // 1 next to the interior
    const bool own = ys == 0 && a.ram && a.fs && j == dn;  // x strip row dn+1 keeps its own history
    float p1v[CPW], c2v[CPW], Av[CPW], Bv[CPW], chx_o[CPW], dphy[CPW], phy_o[CPW], chy_o[CPW], hsc[CPW], hsn[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int i = i0 + w * CPW + q;
        const bool ok = rowok && i < i1;
        p1v[q] = (ok && !a.zp) ? pn[i * ld + j] : 0.f;
        c2v[q] = ok ? c2_at(a, j, i) : 0.f;
        const int distx = xs == 0 ? 0 : (xs == 1 ? dn - i : i - (dn + a.nx) + 1);
        // corners use F(max(dist_x, dist_y)) for both directions (solver_PML_coeffs.f90:288-322)
        const int dist = distx > disty ? distx : disty;
        Av[q] = ok ? fA[dist - 1] : 0.f;
        Bv[q] = ok ? fB[dist - 1] : 0.f;
        chx_o[q] = (ok && xs != 0 && !a.zm) ? gp((const float *)a.chx)[(i - xoff) * ld + j] : 0.f;
        dphy[q] = phy_o[q] = chy_o[q] = 0.f;
        if (ok && ys != 0 && !a.zm) {
            const int jl = ys == 1 ? j : j - (dn + a.ny);
            const int yr = ys == 1 ? j : j - a.ny;  // row of the y memory arrays
            const int yrc = yr - jl + clampi(jl, 3, dn - 4);
            gptr<const float> ph = gp(a.phy_o) + i * dn2;
            float d = (ph[yrc + 1] - ph[yrc - 1]) * kC1_1;
            d = d + (ph[yrc + 2] - ph[yrc - 2]) * kC1_2;
            d = d + (ph[yrc + 3] - ph[yrc - 3]) * kC1_3;
            dphy[q] = d;
            phy_o[q] = ph[yr];
            chy_o[q] = gp((const float *)a.chy)[i * dn2 + yr];
        }
        hsc[q] = (ok && own && !a.zc) ? gp(a.hs_c)[i - xoff] : 0.f;
        hsn[q] = (ok && own && !a.zp) ? gp((const float *)a.hs_n)[i - xoff] : 0.f;
    }
    stage_p<TX>(gp(a.pc), ld, a.nyt, a.nxt, j0, i0, a.zc, tile);
    __syncthreads();
    if (!rowok) return;

#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int i = i0 + w * CPW + q;
        if (i >= i1) continue;
        const int idx = i * ld + j;
        if (a.ram && ys != 0 && xs == 0 && i - dn >= dn) {
            pn[idx] = 0.f;
            continue;
        }
        const float p1 = p1v[q], c2 = c2v[q];
        const int ic = clampi(i, 3, a.nxt - 4);
        // P(gj, gi) = tile[(gi - i0 + 3) * HY + (gj - j0 + 3)]
        const float *Tx = tile + (ic - i0 + 3) * HY + (lane + 3);       // row j, column ic
        const float *Ty = tile + (i - i0 + 3) * HY + (jc - j0 + 3);     // row jc, column i
        const float px = Tx[0], py = Ty[0];
        const float xm1 = Tx[-HY], xp1 = Tx[HY], xm2 = Tx[-2 * HY], xp2 = Tx[2 * HY];
        const float xm3 = Tx[-3 * HY], xp3 = Tx[3 * HY];
        const float ym1 = Ty[-1], yp1 = Ty[1], ym2 = Ty[-2], yp2 = Ty[2];
        const float ym3 = Ty[-3], yp3 = Ty[3];
        const float d2x = ZS ? d2z(px, xp1, xm1, xp2, xm2, xp3, xm3) : d2(px, xp1, xm1, xp2, xm2, xp3, xm3);
        const float d2y = ZS ? d2z(py, yp1, ym1, yp2, ym2, yp3, ym3) : d2(py, yp1, ym1, yp2, ym2, yp3, ym3);
        const float p = tile[(i - i0 + 3) * HY + (lane + 3)];
        const float A = Av[q], B = Bv[q];
        // x-direction memory (strips xl/xr and the corners), pml_matrices_x/_xycorner(1)
        float propx = 0.f;
        if (xs != 0) {
            const int xc = i - xoff;                            // memory column of this cell
            const int xcc = xlo + clampi(xc - xlo, 3, dn - 4);  // clamped derivative column
            const float *ph = phs + (xcc - xm0 + 3) * TY + lane;
            float dphi = (ph[TY] - ph[-TY]) * kC1_1;
            dphi = dphi + (ph[2 * TY] - ph[-2 * TY]) * kC1_2;
            dphi = dphi + (ph[3 * TY] - ph[-3 * TY]) * kC1_3;
            const float pho = phs[(xc - xm0 + 3) * TY + lane];
            const int xo = xc * ld + j;
            const float cho = chx_o[q];
            float d1x = (xp1 - xm1) * kC1_1;
            d1x = d1x + (xp2 - xm2) * kC1_2;
            d1x = d1x + (xp3 - xm3) * kC1_3;
            const float phn = A * pho + B * d1x;
            const float chn = A * cho + B * (d2x + dphi);
            gp(a.phx_n)[xo] = phn;
            gp(a.chx)[xo] = chn;
            propx = dphi + chn;
        }
        float propy = 0.f;
        if (ys != 0) {
            const int yr = ys == 1 ? j : j - a.ny;  // row of the y memory arrays
            const int yo = i * dn2 + yr;
            const float dphi = dphy[q];
            float d1y = (yp1 - ym1) * kC1_1;
            d1y = d1y + (yp2 - ym2) * kC1_2;
            d1y = d1y + (yp3 - ym3) * kC1_3;
            const float phn = A * phy_o[q] + B * d1y;
            const float chn = A * chy_o[q] + B * (d2y + dphi);
            gp(a.phy_n)[yo] = phn;
            gp(a.chy)[yo] = chn;
            propy = dphi + chn;
        }
        float v;
        if (ys == 0) {
            // x strip, interior rows (:376-382).  With a free surface the _ram form
            // zeroes row dn+1 of the grid but the strip keeps its own history there.
            const float ps = own ? hsc[q] : p;
            const float p1s = own ? hsn[q] : p1;
            v = (-p1s + 2.f * ps) + c2 * ((d2x + d2y) + propx);
            if (own) {
                gp(a.hs_n)[i - xoff] = v;
                v = 0.f;
            }
        } else if (xs == 0) {
            v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propy);  // y strip (:388-396)
        } else {
            v = (-p1 + 2.f * p) + c2 * ((d2x + d2y) + propx);  // corner x call (:402-403)
            v = v + c2 * propy;                                 // corner y call (:407)
        }
        pn[idx] = v;
    }
}

// LDS of a step workgroup: a CPML tile's wavefield tile and x-memory staging, or the
// four stream items' snapshot staging (kStageCols columns x 64 lanes x 16 B each).
constexpr int cpml_tile_floats(int TX) { return (TX + 6) * (kTileY + 6); }
constexpr int lds_floats(int TX) {
    return cpml_tile_floats(TX) + (TX + 6) * kTileY > 4 * kStageCols * 64 * 4 ? cpml_tile_floats(TX) + (TX + 6) * kTileY
                                                                              : 4 * kStageCols * 64 * 4;
}

// One work item of one time step of one shot: t < ncpml is a CPML tile (the whole
// workgroup, LDS-staged), otherwise a group of four interior stream items (one per
// wavefront, no workgroup barrier).  The branch is uniform over the workgroup.  SV: the
// snapshot / imaging variant the interior is compiled for (SnapVariant).
template <int TX, int SV>
__device__ __forceinline__ void step_item(const StepArgs &a, int t, float *lds) {
    float *tile = lds, *aux = lds + cpml_tile_floats(TX);
    if (t >= a.ncpml) {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int u = 4 * (t - a.ncpml) + w;
        if (u < a.nstream) stream_item<SV>(a, u, (v4f *)lds + w * kStageCols * 64);
        return;
    }
    int R, C;
    if (t < a.ncpml_a) {  // top / bottom strips
        const int cl = fdiv(t, a.fd_tbr);
        const int r = t - cl * a.tb_rows;
        // both strips' row tiles, or with a free surface the bottom strip's only
        R = a.tb_rows == 2 * a.rt_pml ? (r < a.rt_pml ? r : r + a.rt_int) : r + a.rt_pml + a.rt_int;
        // the listed columns: left corners, the interior column tiles before the skipped
        // ones (_ram pads past column dn, left out), right corners
        C = cl < a.ct_pml + a.ct_int - a.ctb_skip ? cl : cl + a.ctb_skip;
    } else {  // x strips over the interior rows
        const int t2 = t - a.ncpml_a;
        const int cc = fdiv(t2, a.fd_rint);
        R = a.rt_pml + (t2 - cc * a.rt_int);
        C = cc < a.ct_pml ? cc : cc + a.ct_int;
    }
    int j0, j1, i0, i1;
    const int ys = axis_tile(R, a.rt_pml, a.rt_int, a.dn, a.ny, kTileY, a.fd_rpml, &j0, &j1);
    const int xs = axis_tile(C, a.ct_pml, a.ct_int, a.dn, a.nx, TX, a.fd_cpml, &i0, &i1);
    cpml_tile<TX, (SV & kSvZeroSum) != 0>(a, xs, ys, j0, j1, i0, i1, tile, aux);
}

template <int TX, int SV>
__global__ __launch_bounds__(kBlock) void rtm_step_kernel(StepArgs a) {
    __shared__ __attribute__((aligned(16))) float lds[lds_floats(TX)];
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
    if (t >= a.nitems) return;
    step_item<TX, SV>(a, t, lds);  // the one-shot solver: FP32 snapshots, plain imaging
}

// Per-step fields of a slot at global step G (forward steps k = 2..nt, then backward
// steps k = 2..nt; ping-pong buffers alternate with k; each phase starts from zero
// state).  `a` holds the slot's static fields; returns false when the slot is idle.
__device__ __forceinline__ bool resolve_step(const SlotDesc &sd, long long G, StepArgs &a) {
    const long long tl = G - sd.g0;
    const int nt = sd.nt;
    if (sd.g0 < 0 || tl < 0 || tl >= 2LL * (nt - 1)) return false;
    const bool fwd = tl < nt - 1;
    const int k = (int)(fwd ? tl + 2 : tl - (nt - 1) + 2);
    const int m = (k - 2) & 1;
    a.k = k;
    a.mode = fwd ? 0 : 1;
    a.pc = m ? sd.pB : sd.pA;
    a.pn = m ? sd.pA : sd.pB;
    a.phx_o = m ? sd.phxB : sd.phxA;
    a.phx_n = m ? sd.phxA : sd.phxB;
    a.phy_o = m ? sd.phyB : sd.phyA;
    a.phy_n = m ? sd.phyA : sd.phyB;
    a.hs_c = m ? sd.hsB : sd.hsA;
    a.hs_n = m ? sd.hsA : sd.hsB;
    a.zc = k == 2;
    a.zp = k <= 3;
    a.zm = k == 2;
    a.rec_row = -1;
    int slot = -1;
    if (fwd) {
        const int s = fdiv(k - 1, sd.fd_ss);
        if (s * sd.ss == k - 1 && s <= sd.nsnap) slot = s - 1;
        a.ev_ptr = sd.ev_ptr_f;
        a.ev_cell = sd.ev_cell_f;
        a.ev_inj = sd.ev_inj_f;
        a.inj_val = sd.inj_f;
        a.ninj = sd.ninj_f;
    } else {
        // backward step k pairs with the forward snapshot taken at k_f = nt + 2 - k
        const int s = fdiv(nt + 1 - k, sd.fd_ss);
        if (s * sd.ss == nt + 1 - k && s >= 1 && s <= sd.nsnap) slot = s - 1;
        a.ev_ptr = sd.ev_ptr_b;
        a.ev_cell = sd.ev_cell_b;
        a.ev_inj = sd.ev_inj_b;
        a.inj_val = sd.inj_b;
        a.ninj = sd.ninj_b;
    }
    a.snap_slot = slot;
    return true;
}

// One time step of every active slot of a launch group: a grid of workgroups walks the
// group's compact work list (the occupied slots' work items back to back, so idle slots
// and narrow windows launch nothing); each XCD takes a contiguous range of items
// (neighbouring tiles of one shot).
// The slot descriptors are read-only during a launch (the host rewrites them between
// chunks), so their words are fetched with scalar loads.
// CMP = BatchArgs compact (a separate instantiation, so that the full-slot mapping reads
// nothing but its launch arguments before the first item's descriptor)
template <int TX, int SV, bool CMP>
__global__ __launch_bounds__(kBlock) void rtm_batch_kernel(BatchArgs b, const SlotDesc *__restrict__ slots,
                                                           const long long *__restrict__ gbase) {
    __shared__ __attribute__((aligned(16))) float lds[lds_floats(TX)];
    const long long G = *gbase + b.goff;
    const int n = b.nslots;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    // the work list: with compact = 0 (every slot full) cap_tiles items per slot, read from
    // the launch arguments alone; else the group's compact list, the items of its occupied
    // slots back to back
    int total = n * b.cap_tiles, uni = b.cap_tiles;
    FastDiv wfd = b.fd_cap;
    // groups of <= 64 slots: each slot's first item in the compact list, in LDS (read per item
    // by the lane of the slot's index; kept out of registers for the step's register budget)
    __shared__ int woffs[64];
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
        // the slot holding item wk: wk / uni when every slot holds uni items, else the last
        // slot whose range starts at or before wk -- with <= 64 slots one wave-wide compare
        // of the lanes' first items (the ranges are in slot order), else a walk from the
        // guess wk / (mean items per slot)
        int s, t;
        if (!CMP || uni > 0) {
            s = fdiv(wk, wfd);
            t = wk - s * uni;
        } else if (n <= 64) {
            const int mine = woffs[threadIdx.x & 63];
            s = __popcll(__ballot(mine <= wk)) - 1;
            t = wk - __builtin_amdgcn_readlane(mine, s);
        } else {
            s = min(fdiv(wk, wfd), n - 1);
            while (s > 0 && slots[s].woff > wk) --s;
            while (s + 1 < n && slots[s + 1].woff <= wk) ++s;
            t = wk - slots[s].woff;
        }
        const SlotDesc &sd = slots[s];
        StepArgs a = sd.a;
        if (!resolve_step(sd, G, a) || t >= a.nitems) continue;
        step_item<TX, SV>(a, t, lds);
        __syncthreads();  // the LDS buffers are reused by the next item
    }
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
__global__ void rtm_stack_kernel(float *img, float *comp, int ld_img, int ny, int nx_w, int x0, double *sum,
                                 int ld_sum, int *fold) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j < ny && i < nx_w) {
        const size_t o = (size_t)i * ld_img + j;
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

// taper_pad_trace's half-cosine taper (filters.f90:20-48) on resident snapshots: slot
// s0 + q (0-based, q < nq) of every interior cell is multiplied by w[q] (FP32, as the
// driver's trace(i0+n) * w on Fields, calculate_PSDM_RTM.f90:421-427).
__global__ void rtm_taper_snaps_kernel(float *snaps, size_t stride, int lds, int ny, int nx, int s0, const float *w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, q = blockIdx.z;
    if (j < ny && i < nx) {
        float *f = snaps + (size_t)(s0 + q) * stride + (size_t)i * lds + j;
        *f = *f * w[q];
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
template <int TX>
void launch_step_tx(const StepArgs &a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    auto k = a.zsum ? rtm_step_kernel<TX, kSvF32 + kSvZeroSum> : rtm_step_kernel<TX, kSvF32>;
    hipExtLaunchKernelGGL(k, dim3(8 * a.per_xcd), dim3(kBlock), 0, s, e0, e1, 0, a);
}

void launch_step(const StepArgs &a, int tile_x, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (tile_x == 32)
        launch_step_tx<32>(a, s, e0, e1);
    else if (tile_x == 16)
        launch_step_tx<16>(a, s, e0, e1);
    else
        launch_step_tx<8>(a, s, e0, e1);
}

template <int TX, int SV>
void launch_batch_tx(const BatchArgs &b, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    auto k = b.compact ? rtm_batch_kernel<TX, SV, true> : rtm_batch_kernel<TX, SV, false>;
    if (!e0 && !e1)  // plain launch: also valid under stream capture
        hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, b, b.slots, b.gbase);
    else
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0, b, b.slots, b.gbase);
}

template <int SV>
void launch_batch_sv(const BatchArgs &b, int tile_x, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (tile_x == 32)
        launch_batch_tx<32, SV>(b, grid, s, e0, e1);
    else if (tile_x == 16)
        launch_batch_tx<16, SV>(b, grid, s, e0, e1);
    else
        launch_batch_tx<8, SV>(b, grid, s, e0, e1);
}

void launch_batch(const BatchArgs &b, int tile_x, int sv, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    grid = (grid + 7) / 8 * 8;
    constexpr int Z = kSvZeroSum;
    switch (sv) {
    case kSvF32Kahan: launch_batch_sv<kSvF32Kahan>(b, tile_x, grid, s, e0, e1); break;
    case kSv16: launch_batch_sv<kSv16>(b, tile_x, grid, s, e0, e1); break;
    case kSv16Kahan: launch_batch_sv<kSv16Kahan>(b, tile_x, grid, s, e0, e1); break;
    case kSvF32 + Z: launch_batch_sv<kSvF32 + Z>(b, tile_x, grid, s, e0, e1); break;
    case kSvF32Kahan + Z: launch_batch_sv<kSvF32Kahan + Z>(b, tile_x, grid, s, e0, e1); break;
    case kSv16 + Z: launch_batch_sv<kSv16 + Z>(b, tile_x, grid, s, e0, e1); break;
    case kSv16Kahan + Z: launch_batch_sv<kSv16Kahan + Z>(b, tile_x, grid, s, e0, e1); break;
    default: launch_batch_sv<kSvF32>(b, tile_x, grid, s, e0, e1); break;
    }
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

void launch_stack(float *img, float *comp, int ld_img, int ny, int nx_w, int x0, double *sum, int ld_sum, int *fold,
                  hipStream_t s) {
    hipLaunchKernelGGL(rtm_stack_kernel, dim3((ny + 255) / 256, nx_w), dim3(256), 0, s, img, comp, ld_img, ny, nx_w, x0,
                       sum, ld_sum, fold);
}

void launch_taper_snaps(float *snaps, size_t stride, int lds, int ny, int nx, int s0, int nq, const float *w,
                        hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(rtm_taper_snaps_kernel, dim3((ny + 255) / 256, nx, nq), dim3(256), 0, s, snaps, stride, lds, ny,
                       nx, s0, w);
}

void launch_normalize(const double *sum, const int *fold, float *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(rtm_normalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sum, fold, out, n);
}

}  // namespace psdm
