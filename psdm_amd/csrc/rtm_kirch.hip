// Kirchhoff depth migration of one shot on the device: eikonal travel times by fast
// sweeping and the diffraction stack with the aperture-angle weight.
//
//   eikonal_fsm_kernel     solve_eikonal_FSM (solve_eikonal_FSM.f90:1-83) for one source
//                          point per workgroup, bit for bit: four Gauss-Seidel sweeps, each
//                          run as a wavefront over the sweep's diagonals.  In a sweep the
//                          cell at sweep-order (row r, column c) reads the updated values of
//                          (r-1, c) and (r, c-1) and the previous sweep's values of (r+1, c)
//                          and (r, c+1), so the cells of diagonal d = r + c depend only on
//                          diagonal d-1 (new) and d+1 (old): processing the diagonals in
//                          order with a barrier between them gives the sequential sweep's
//                          results exactly.  Each thread owns model columns and walks down
//                          them one row per step (loads stream down the column); the two
//                          values a column's neighbours need each step -- its new value and
//                          its not-yet-updated next row -- pass through LDS (double
//                          buffered, one barrier per step).
//   kirchhoff_image_kernel the imaging loop of calculate_PSDM_kirchoff.f90:394-447: per
//                          pixel the sum over receivers (in receiver order) of the weight
//                          times the central time derivative of the trace at the total
//                          travel time.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rtm_ctx.hpp"

using psdm::host::fail;
using psdm::host::get_buf;
using psdm::host::set_device;

namespace psdm {
namespace {

constexpr float kBig = 1.0e9f;  // initial travel time (BIG, solve_eikonal_FSM.f90:11)
constexpr int kMaxCols = 8;     // model columns per thread (nx <= 8 * 1024)
constexpr int kMaxBlock = 1024;  // eikonal workgroup size limit (its __launch_bounds__)
constexpr int kTrash = 1024;     // per-solve sink for the stores of lanes without a cell, one word per thread
static_assert(kTrash >= kMaxBlock, "every thread of a workgroup needs its own trash word");

// Division by a divisor fixed for the launch (local_solver's dx*dx, dy*dy and 2*aa).
// With zh = RN(1/y) and zl = RN(1/y - zh), RN(x*zh + RN(x*zl)) -- one product and one
// FMA -- is the correctly rounded x / y for every x when a condition on y holds
// (Brisebarre, Muller and Raina, "Accelerating correctly rounded floating-point division
// when the divisor is known in advance", IEEE Trans. Computers 53(8), 2004).  The host
// establishes it per divisor by comparing both forms on the device for every significand
// (x in [1, 2), divcheck_kernel); it then holds for every |x| in [2^-60, 2^96] and for
// 0, because scaling x by a power of two scales every rounding of the form alike while no
// product underflows or overflows (the host enables it only for 2^-30 <= y <= 2^30 and
// zl = 0 or |zl| >= 2^-60, so |x*zl| >= 2^-120 and |x*zh| <= 2^126).  local_solver's
// numerators are a, b, a*a, b*b (a, b <= BIG = 1e9: travel times only ever decrease from
// it) and -bb + sqrt(disc) >= 2*(a/(dx*dx) + b/(dy*dy)): all of them are 0 or lie in that
// range when a and b are each 0 or >= 2^-30, checked once per cell, and the last one is
// <= 2^96 (checked: a near-zero velocity can make it larger, or infinite); any other
// numerator, and a divisor that fails the check, take the IEEE division.
struct EikConsts {
    float dx, dy, aa;     // aa = 1/(dx*dx) + 1/(dy*dy) in the reference's order
    float zh[3], zl[3];   // per divisor dx*dx, dy*dy, 2*aa
    int fast;             // bit i: divisor i takes the two-operation form
};

template <bool kAllow>
__device__ __forceinline__ float divc(float x, float y, float zh, float zl, int fast, int i, bool in_range = true) {
    if (kAllow && ((fast >> i) & 1) && in_range) return __builtin_fmaf(x, zh, x * zl);
    return x / y;
}

// local_solver (solve_eikonal_FSM.f90:63-79), the reference's FP32 operations in order
// (every quotient correctly rounded, by divc)
template <bool kAllow>
__device__ __forceinline__ float local_solver_q(float a, float b, const EikConsts &k, float v_inv) {
    const float d0 = k.dx * k.dx, d1 = k.dy * k.dy, aa = k.aa;
    const float bb = -2.0f * (divc<kAllow>(a, d0, k.zh[0], k.zl[0], k.fast, 0) +
                              divc<kAllow>(b, d1, k.zh[1], k.zl[1], k.fast, 1));
    const float cc = divc<kAllow>(a * a, d0, k.zh[0], k.zl[0], k.fast, 0) +
                     divc<kAllow>(b * b, d1, k.zh[1], k.zl[1], k.fast, 1) - v_inv * v_inv;
    const float disc = bb * bb - 4.0f * aa * cc;
    const float dmin = fminf(k.dx, k.dy);
    float t;
    if (disc >= 0.0f) {
        const float x3 = -bb + sqrtf(disc);
        t = divc<kAllow>(x3, 2.0f * aa, k.zh[2], k.zl[2], k.fast, 2, x3 <= 0x1p96f);
        if (t < fmaxf(a, b)) t = fmaxf(a, b) + v_inv * dmin;
    } else {
        t = fminf(a, b) + v_inv * dmin;
    }
    return t;
}

__device__ __forceinline__ float local_solver(float a, float b, const EikConsts &k, float v_inv) {
    const bool ok = (a == 0.0f || a >= 0x1p-30f) && (b == 0.0f || b >= 0x1p-30f);
    return ok ? local_solver_q<true>(a, b, k, v_inv) : local_solver_q<false>(a, b, k, v_inv);
}

// Three divisors fixed for a launch and their two-operation constants (divc).
struct DivSet {
    float y[3], zh[3], zl[3];
};

// Per divisor of `k`: whether RN(x*zh + RN(x*zl)) equals x / y for every x in [1, 2).
// Thread t of block b tests the 32 significands from 32 * (b * blockDim + t); bad[3b + i]
// = 1 when block b found a mismatch for divisor i.
constexpr int kDivCheckPer = 32, kDivCheckBlock = 256;
constexpr int kDivCheckBlocks = (1 << 23) / (kDivCheckPer * kDivCheckBlock);
__global__ __launch_bounds__(kDivCheckBlock) void divcheck_kernel(DivSet k, int *bad) {
    const unsigned base = (blockIdx.x * blockDim.x + threadIdx.x) * (unsigned)kDivCheckPer;
    int miss[3] = {0, 0, 0};
    for (int q = 0; q < kDivCheckPer; ++q) {
        const float x = __uint_as_float(0x3f800000u | (base + (unsigned)q));
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float f = __builtin_fmaf(x, k.zh[i], x * k.zl[i]);
            if (__float_as_uint(f) != __float_as_uint(x / k.y[i])) miss[i] = 1;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int m = __syncthreads_or(miss[i]);
        if (threadIdx.x == 0) bad[blockIdx.x * 3 + i] = m;
    }
}

// Whether x is 0 or in [2^-60, 2^96], the range over which a divisor that passed the check
// gives correctly rounded quotients (see divc).
__device__ __forceinline__ bool div_range(float x) {
    const float ax = fabsf(x);
    return x == 0.0f || (ax >= 0x1p-60f && ax <= 0x1p96f);
}

// Workgroup barrier for the wavefront steps: LDS writes complete, then s_barrier.  Global
// memory is not fenced (that would wait for the prefetch loads every step): within a sweep
// each thread reads and writes only its own columns; the sweeps are separated by full
// __syncthreads().
__device__ __forceinline__ void lds_barrier() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// 1/v once per cell (v_inv = 1.0/v(j,i), solve_eikonal_FSM.f90:52, the same IEEE division),
// stored in the two diagonal-major layouts the sweeps read (see eikonal_fsm_kernel).
__global__ void reciprocal_kernel(const float *__restrict__ v, int ny, int nx, float *__restrict__ va,
                                  float *__restrict__ vm) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nx * ny) return;
    const int i = (int)(e / ny), j = (int)(e - (size_t)i * ny);
    const float r = 1.0f / v[e];
    va[(size_t)(i + j) * nx + i] = r;
    vm[(size_t)(nx - 1 - i + j) * nx + (nx - 1 - i)] = r;
}

// Travel-time layouts of one solve: A[(i + j) * nx + i] (anti-diagonals contiguous), M[(nx-1-i
// + j) * nx + (nx-1-i)] (diagonals contiguous) and the model's column-major i * ny + j.
enum { kLayA = 0, kLayM = 1, kLayModel = 2 };
constexpr int kTile = 64, kTileLd = kTile + 1;
__device__ __forceinline__ size_t lay_addr(int kind, int i, int j, int nx, int ny) {
    return kind == kLayA ? (size_t)(i + j) * nx + i
                         : kind == kLayM ? (size_t)(nx - 1 - i + j) * nx + (nx - 1 - i) : (size_t)i * ny + j;
}
// Slot s of a 64 x 64 tile, numbered so that consecutive slots are consecutive addresses of
// the layout: (il, jl), false for slots outside the tile.
__device__ __forceinline__ bool tile_slot(int kind, int s, int &il, int &jl) {
    if (kind == kLayModel) {
        il = s >> 6;
        jl = s & 63;
        return true;
    }
    const int e = s >> 6, f = s & 63;  // e: diagonal within the tile, f: position along it
    il = kind == kLayA ? f : kTile - 1 - f;
    jl = e - f;
    return jl >= 0 && jl < kTile;
}
// Copies a solve's travel times between layouts through 64 x 64 LDS tiles (both sides
// coalesced).  The caller separates it from the accesses around it with __syncthreads().
__device__ void relayout(const float *__restrict__ src, int sk, float *__restrict__ dst, int dk, int nx, int ny,
                         float *tile) {
    const int ns = sk == kLayModel ? kTile * kTile : (2 * kTile - 1) * kTile;
    const int nd = dk == kLayModel ? kTile * kTile : (2 * kTile - 1) * kTile;
    for (int i0 = 0; i0 < nx; i0 += kTile)
        for (int j0 = 0; j0 < ny; j0 += kTile) {
            for (int q = threadIdx.x; q < ns; q += blockDim.x) {
                int il, jl;
                if (!tile_slot(sk, q, il, jl) || i0 + il >= nx || j0 + jl >= ny) continue;
                tile[il * kTileLd + jl] = src[lay_addr(sk, i0 + il, j0 + jl, nx, ny)];
            }
            lds_barrier();
            for (int q = threadIdx.x; q < nd; q += blockDim.x) {
                int il, jl;
                if (!tile_slot(dk, q, il, jl) || i0 + il >= nx || j0 + jl >= ny) continue;
                dst[lay_addr(dk, i0 + il, j0 + jl, nx, ny)] = tile[il * kTileLd + jl];
            }
            lds_barrier();
        }
}

// Offset in a sweep's layout of the cell of step (diagonal) D, sweep-order column c.
__device__ __forceinline__ unsigned ring_addr(bool fwd, int D, int c, int nx, int dmax) {
    return fwd ? (unsigned)(D * nx + c) : (unsigned)((dmax - D) * nx + (nx - 1 - c));
}

// K model columns per thread; P: prefetch depth, in rows, of each column's old travel times
// and 1/v.
//
// A sweep processes the cells of one diagonal per step, so the travel times are kept in the
// layout whose diagonals are contiguous for that sweep -- A for sweeps 1 and 3
// (anti-diagonals), M for sweeps 2 and 4 -- and every load and store of a step coalesces.
// A sweep reads and writes its layout in place (a step never reads a cell that an earlier
// step of the same sweep wrote: its new neighbours come through LDS and registers); between
// sweeps the travel times are re-laid out through LDS tiles (A -> M -> A -> M), and after
// the last sweep M -> the model's column-major output.
template <int K, int P>
__global__ __launch_bounds__(kMaxBlock) void eikonal_fsm_kernel(const float *__restrict__ va, const float *__restrict__ vm,
                                                           int ny, int nx, EikConsts kc,
                                                           const int *__restrict__ src_x,
                                                           const int *__restrict__ src_z, float *__restrict__ work,
                                                           float *__restrict__ tt) {
    extern __shared__ float lds[];
    float *nw = lds, *od = lds + 2 * nx;  // [2][nx] each: new value at the current row, old value of the next row
    float *tile = lds + 4 * nx;           // relayout tile, 64 x 65
    const int B = blockDim.x, tid = threadIdx.x;
    const size_t nd = (size_t)(nx + ny - 1) * nx;  // diagonal-major array (padding cells never read)
    float *TA = work + (size_t)blockIdx.x * (2 * nd + kTrash), *TM = TA + nd;  // then kTrash words
    float *T = tt + (size_t)blockIdx.x * nx * ny;
    for (size_t e = tid; e < nd; e += B) TA[e] = kBig;
    __syncthreads();
    if (tid == 0) {
        const int i = src_x[blockIdx.x] - 1, j = src_z[blockIdx.x] - 1;
        TA[(size_t)(i + j) * nx + i] = 0.0f;
    }
    __syncthreads();
    constexpr int R = P + 2;  // ring: rows r .. r + P + 1 of the column
    for (int sweep = 1; sweep <= 4; ++sweep) {
        // sweep_pass (:24-58): 1 i up j up, 2 i down j up, 3 i down j down, 4 i up j down.
        // Column c (sweep order) is at row r = d - c in step d.  Per column the thread keeps
        // the old T and 1/v of rows r .. r+P+1 in a ring (tq, vq; the old T(r+1) is what the
        // previous column needs from it this step) and prv = new T(r-1).
        // sweep-order cell (r, c) of step d lies at D * nx + X of the sweep's layout, with
        // D = d, X = c (sweeps 1, 2) or D = nx+ny-2-d, X = nx-1-c (sweeps 3, 4)
        const bool fwd = sweep <= 2;
        float *Ts = (sweep == 1 || sweep == 3) ? TA : TM;
        const float *Vs = (sweep == 1 || sweep == 3) ? va : vm;
        const unsigned sink = (unsigned)(Ts == TA ? 2 * nd : nd);  // the trash words, from Ts
        const int dmax = nx + ny - 2;
        float tq[K][R], vq[K][R], prv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int c = tid + k * B;
            prv[k] = 0.f;
#pragma unroll
            for (int m = 0; m < R; ++m) {  // slot m: diagonal m (row m - c)
                tq[k][m] = vq[k][m] = 0.f;
                if (K <= 2 || (c < nx && m - c >= 0 && m - c < ny)) {
                    const unsigned q = K <= 2 ? ring_addr(fwd, min(m, dmax), min(c, nx - 1), nx, dmax)
                                              : ring_addr(fwd, m, c, nx, dmax);
                    tq[k][m] = Ts[q];
                    vq[k][m] = Vs[q];
                }
            }
        }
        const int nsteps = nx + ny - 1;
        // One wavefront step: diagonal d, ring sub-step u (row r + m of a column sits in slot
        // (u + m) % R, so no register of an in-flight load is ever moved).
        auto step = [&](const int d, const int u) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int c = tid + k * B;
                if (c < nx) od[(d & 1) * nx + c] = tq[k][(u + 1) % R];
            }
            lds_barrier();
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int c = tid + k * B, r = d - c;
                if constexpr (K > 2) {
                    if (c >= nx) continue;
                }
                const bool live = c < nx && r >= 0 && r < ny;
                float tnew = 0.0f;
                if (live) {
                    const float self = tq[k][u];
                    // neighbours in sweep order: previous column (new, LDS), next column
                    // (old, its lookahead in LDS), previous row (new, register), next row (old)
                    const float cp = c > 0 ? nw[((d - 1) & 1) * nx + c - 1] : self;
                    const float cn = c + 1 < nx ? od[(d & 1) * nx + c + 1] : self;
                    const float rp = r > 0 ? prv[k] : self;
                    const float rn = r + 1 < ny ? tq[k][(u + 1) % R] : self;
                    tnew = self;
                    if (self != 0.0f) {  // the source point keeps 0
                        // a = min(T(j,i-1), T(j,i+1)), b = min(T(j-1,i), T(j+1,i)): min is
                        // symmetric, so the sweep-order neighbours give the same values
                        const float a = fminf(cp, cn), b = fminf(rp, rn);
                        const float t = local_solver(a, b, kc, vq[k][u]);
                        if (t < self) tnew = t;
                    }
                    prv[k] = tnew;
                    nw[(d & 1) * nx + c] = tnew;
                }
                // slot u now takes row r + R (diagonal d + R)
                if constexpr (K <= 2) {
                    // Every lane stores and reloads, outside any branch: with the same memory
                    // instructions on every path the compiler's vmcnt waits count back through
                    // the ring (after a lane-conditional load it assumes none were issued
                    // since, and drains the queue every step).  A lane without a cell stores
                    // into its trash word; the reload is clamped into the array (a row
                    // outside the grid is never consumed).  Wider rings (K >= 4) would spill
                    // with the extra address registers and keep the conditional form.
                    const int cl = min(c, nx - 1);
                    Ts[live ? ring_addr(fwd, d, cl, nx, dmax) : sink + (unsigned)tid] = tnew;
                    const unsigned q = ring_addr(fwd, min(d + R, dmax), cl, nx, dmax);
                    tq[k][u] = Ts[q];
                    vq[k][u] = Vs[q];
                } else {
                    if (live) Ts[ring_addr(fwd, d, c, nx, dmax)] = tnew;
                    const int rl = r + R;
                    if (c < nx && rl >= 0 && rl < ny) {
                        const unsigned q = ring_addr(fwd, d + R, c, nx, dmax);
                        tq[k][u] = Ts[q];
                        vq[k][u] = Vs[q];
                    }
                }
            }
        };
        if constexpr (K <= 2) {
            // Whole blocks of R steps, then the tail: a uniform early exit inside the unrolled
            // block would join the loop's back edge with paths that issued fewer loads, and
            // the vmcnt waits of the next block's first steps would drain the ring again.
            const int nfull = nsteps - nsteps % R;
            for (int d0 = 0; d0 < nfull; d0 += R) {
#pragma unroll
                for (int u = 0; u < R; ++u) step(d0 + u, u);
            }
#pragma unroll
            for (int u = 0; u < R - 1; ++u)
                if (nfull + u < nsteps) step(nfull + u, u);
        } else {
            for (int d0 = 0; d0 < nsteps; d0 += R) {
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    if (d0 + u >= nsteps) break;  // uniform over the workgroup
                    step(d0 + u, u);
                }
            }
        }
        __syncthreads();
        if (sweep == 1 || sweep == 3)
            relayout(TA, kLayA, TM, kLayM, nx, ny, tile);
        else if (sweep == 2)
            relayout(TM, kLayM, TA, kLayA, nx, ny, tile);
        else
            relayout(TM, kLayM, T, kLayModel, nx, ny, tile);
        __syncthreads();
    }
}

// The imaging loop's launch constants beside its arguments: the two-operation division
// (divc) by dt and by 2*dt where it passed the check (bit 0 / bit 1 of `fast`), and
// cos_in = cos(theta_max) + 1e-4: a pair with cos_theta >= cos_in has theta below
// theta_max by at least 1e-4 / sin(theta_max) radians, far beyond the error of acosf and
// of the degree conversion, so its aperture weight is 1 without evaluating acosf.  Both
// leave every result as the IEEE form gives it.
struct ImgConsts {
    float zh_t, zl_t, zh_2t, zl_2t;
    int fast;
    float cos_in;
};

__global__ __launch_bounds__(256) void kirchhoff_image_kernel(const float *__restrict__ tts, const float *__restrict__ ttr,
                                                              const float *__restrict__ data, int ld, int nrec,
                                                              int nt_prop, int ny, int nx, float dd, float dt,
                                                              float vp0, float theta_max, float theta_taper,
                                                              float wmax, ImgConsts kc, float *__restrict__ img) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;  // pixel (j, i), j fastest
    if (p >= nx * ny) return;
    const int j = p % ny;
    const float tmin = 0.02f, eps = 1.0e-6f, pi = 3.14159265f;  // calculate_PSDM_kirchoff.f90:8-10
    const float z_img = (float)j * dd;
    const float ts = tts[p];
    const float rs = vp0 * fmaxf(ts, tmin);
    const size_t n = (size_t)nx * ny;
    float acc = 0.0f;
    for (int k = 0; k < nrec; ++k) {
        const float tr = ttr[(size_t)k * n + p];
        const float tt_tot = ts + tr;
        const float q = (kc.fast & 1) && div_range(tt_tot) ? __builtin_fmaf(tt_tot, kc.zh_t, tt_tot * kc.zl_t)
                                                          : tt_tot / dt;
        const int it = 1 + (int)floorf(q);
        if (it <= 1 || it >= nt_prop) continue;
        const float rr = vp0 * fmaxf(tr, tmin);
        float cos_theta = z_img / (rs + rr + 1.e-6f);
        cos_theta = fmaxf(fminf(cos_theta, 1.0f), 0.0f);
        float w_theta = 1.0f;
        if (!(cos_theta >= kc.cos_in)) {
            const float theta = acosf(cos_theta) * 180.0f / pi;
            if (!(theta <= theta_max)) {
                const float u = (theta - theta_max) / theta_taper;
                w_theta = expf(-(u * u));
            }
        }
        const float weight = w_theta * fminf(cos_theta / sqrtf(rs * rr + eps), wmax);
        const float *trace = data + (size_t)k * ld;
        const float d2 = trace[it] - trace[it - 2];  // Data_target(it+1) - Data_target(it-1)
        const float amp = (kc.fast & 2) && div_range(d2) ? __builtin_fmaf(d2, kc.zh_2t, d2 * kc.zl_2t) : d2 / (2.0f * dt);
        acc = acc + weight * amp;
    }
    img[p] = acc;
}

}  // namespace
}  // namespace psdm

using namespace psdm;

namespace {

int check_eikonal(psdm_rtm_ctx *c, int ny, int nx, float dx, float dy, int nsrc, const int *src_x, const int *src_z) {
    if (!src_x || !src_z) return fail(c, PSDM_RTM_EINVAL, "eikonal: null argument");
    if (ny < 1 || nx < 1 || nx > kMaxCols * kMaxBlock || nsrc < 1)
        return fail(c, PSDM_RTM_EINVAL, "eikonal: need ny >= 1, 1 <= nx <= 8192, nsrc >= 1");
    if (!(dx > 0.f) || !(dy > 0.f)) return fail(c, PSDM_RTM_EINVAL, "eikonal: dx, dy must be > 0");
    if ((long long)(nx + ny - 1) * nx > 2147483647LL - 1024)
        return fail(c, PSDM_RTM_EINVAL, "eikonal: (nx + ny - 1) * nx must stay below 2^31");
    for (int s = 0; s < nsrc; ++s)
        if (src_x[s] < 1 || src_x[s] > nx || src_z[s] < 1 || src_z[s] > ny)
            return fail(c, PSDM_RTM_EINVAL, "eikonal: source point outside the model (1-based indices)");
    return 0;
}

// Sources solved per launch: as many as the workspace limit holds, at least one.
int eikonal_batch(const psdm_rtm_ctx *c, int ny, int nx, int nsrc) {
    const size_t per = (2 * (size_t)(nx + ny - 1) * nx + kTrash) * sizeof(float);
    const size_t fit = c->eik_work_cap / per;
    return (int)std::max<size_t>(1, std::min<size_t>(fit, (size_t)nsrc));
}

// Uploads the source points, computes 1/v of the model extended below and grows the
// workspace to nb solves.  Returns the device source list and the workspace.
int eikonal_setup(psdm_rtm_ctx *c, const float *d_vel, int ny, int nx, int nsrc, const int *src_x, const int *src_z,
                  int nb, const int **sx, const int **sz, const float **va, const float **vm, float **work) {
    void *d_src, *d_v, *d_work;
    const size_t nd = (size_t)(nx + ny - 1) * nx;
    if (int rc = get_buf(c, "eik_src", 2 * (size_t)nsrc * sizeof(int), &d_src)) return rc;
    if (int rc = get_buf(c, "eik_vinv", 2 * nd * sizeof(float), &d_v)) return rc;
    if (int rc = get_buf(c, "eik_work", (2 * nd + kTrash) * (size_t)nb * sizeof(float), &d_work)) return rc;
    std::vector<int> h(2 * (size_t)nsrc);
    std::memcpy(h.data(), src_x, nsrc * sizeof(int));
    std::memcpy(h.data() + nsrc, src_z, nsrc * sizeof(int));
    HIP_TRY(c, hipMemcpyAsync(d_src, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // the host list goes out of scope at return
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *sx = (const int *)d_src;
    *sz = *sx + nsrc;
    float *a = (float *)d_v;
    const size_t ncell = (size_t)nx * ny;
    hipLaunchKernelGGL(reciprocal_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, c->stream, d_vel, ny, nx,
                       a, a + nd);
    HIP_TRY(c, hipGetLastError());
    *va = a;
    *vm = a + nd;
    *work = (float *)d_work;
    return 0;
}

// zh, zl of up to three divisors and the bit mask of those that take the two-operation
// form of divc: divisors in [2^-30, 2^30] with zl = 0 or |zl| >= 2^-60 whose form matched
// the IEEE division for every significand on the device (divcheck_kernel, run once per
// divisor value and kept by the ctx).
int fast_divisors(psdm_rtm_ctx *c, const float ys[3], DivSet *k, int *mask_out) {
    int mask = 0, todo = 0;
    for (int i = 0; i < 3; ++i) {
        const float zh = 1.0f / ys[i];
        const float zl = (float)(1.0 / (double)ys[i] - (double)zh);
        k->y[i] = ys[i];
        k->zh[i] = zh;
        k->zl[i] = zl;
        if (!(ys[i] >= 0x1p-30f && ys[i] <= 0x1p30f && (zl == 0.0f || std::fabs(zl) >= 0x1p-60f))) continue;
        unsigned bits;
        std::memcpy(&bits, &ys[i], 4);
        auto it = c->div_ok.find(bits);
        if (it == c->div_ok.end())
            todo |= 1 << i;
        else if (it->second)
            mask |= 1 << i;
    }
    if (todo) {
        void *v;
        if (int rc = get_buf(c, "divchk", (size_t)kDivCheckBlocks * 3 * sizeof(int), &v)) return rc;
        hipLaunchKernelGGL(divcheck_kernel, dim3(kDivCheckBlocks), dim3(kDivCheckBlock), 0, c->stream, *k, (int *)v);
        HIP_TRY(c, hipGetLastError());
        std::vector<int> bad((size_t)kDivCheckBlocks * 3);
        HIP_TRY(c, hipMemcpyAsync(bad.data(), v, bad.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        int fail_bits = 0;
        for (size_t q = 0; q < bad.size(); ++q)
            if (bad[q]) fail_bits |= 1 << (int)(q % 3);
        for (int i = 0; i < 3; ++i) {
            if (!((todo >> i) & 1)) continue;
            unsigned bits;
            std::memcpy(&bits, &ys[i], 4);
            const bool ok = !((fail_bits >> i) & 1);
            c->div_ok[bits] = ok;
            if (ok) mask |= 1 << i;
        }
    }
    *mask_out = mask;
    return 0;
}

// local_solver's constants for the grid spacings: aa = 1/(dx*dx) + 1/(dy*dy) in the
// reference's order, and per divisor (dx*dx, dy*dy, 2*aa) zh, zl and whether divc may use
// them.
int eikonal_consts(psdm_rtm_ctx *c, float dx, float dy, EikConsts *k, int *fast_mask) {
    const float d0 = dx * dx, d1 = dy * dy;
    const float aa = 1.0f / d0 + 1.0f / d1;
    k->dx = dx;
    k->dy = dy;
    k->aa = aa;
    const float ys[3] = {d0, d1, 2.0f * aa};
    DivSet ds;
    int mask;
    if (int rc = fast_divisors(c, ys, &ds, &mask)) return rc;
    for (int i = 0; i < 3; ++i) {
        k->zh[i] = ds.zh[i];
        k->zl[i] = ds.zl[i];
    }
    k->fast = mask;
    if (fast_mask) *fast_mask = mask;
    return 0;
}

// One launch: sources sx[0..n) into tt (ny, nx, n); e0 / e1 stamp its start / end (or null).
void eikonal_launch(psdm_rtm_ctx *c, const float *va, const float *vm, int ny, int nx, const EikConsts &kc, int n,
                    const int *sx, const int *sz, float *work, float *tt, hipEvent_t e0, hipEvent_t e1) {
    const int B = nx >= kMaxBlock ? kMaxBlock : (nx + 63) / 64 * 64;
    const int K = (nx + B - 1) / B;
    const unsigned lds = (4u * (unsigned)nx + (unsigned)(kTile * kTileLd)) * (unsigned)sizeof(float);
    if (K == 1)
        hipExtLaunchKernelGGL(eikonal_fsm_kernel<1, 4>, dim3(n), dim3(B), lds, c->stream, e0, e1, 0, va, vm, ny, nx, kc,
                              sx, sz, work, tt);
    else if (K == 2)
        hipExtLaunchKernelGGL(eikonal_fsm_kernel<2, 4>, dim3(n), dim3(B), lds, c->stream, e0, e1, 0, va, vm, ny, nx, kc,
                              sx, sz, work, tt);
    else if (K <= 4)
        hipExtLaunchKernelGGL(eikonal_fsm_kernel<4, 2>, dim3(n), dim3(B), lds, c->stream, e0, e1, 0, va, vm, ny, nx, kc,
                              sx, sz, work, tt);
    else
        hipExtLaunchKernelGGL(eikonal_fsm_kernel<8, 1>, dim3(n), dim3(B), lds, c->stream, e0, e1, 0, va, vm, ny, nx, kc,
                              sx, sz, work, tt);
}

int eikonal_events(psdm_rtm_ctx *c) {
    for (int q = 0; q < 2; ++q)
        if (!c->kir_ev[q]) HIP_TRY(c, hipEventCreate(&c->kir_ev[q]));
    return 0;
}

// Adds the stamped launch's device time to the eikonal total; `call` counts one call.
int eikonal_time(psdm_rtm_ctx *c, bool call) {
    float ms;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->kir_ev[0], c->kir_ev[1]));
    c->kir_ms[0] += ms;
    if (call) c->kir_n[0]++;
    return 0;
}

}  // namespace

extern "C" int psdm_rtm_eikonal(psdm_rtm_ctx *c, const float *d_vel, int ny, int nx, float dx, float dy, int nsrc,
                                const int *src_x, const int *src_z, float *d_tt) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_vel || !d_tt) return fail(c, PSDM_RTM_EINVAL, "eikonal: null argument");
    if (int rc = check_eikonal(c, ny, nx, dx, dy, nsrc, src_x, src_z)) return rc;
    if (int rc = set_device(c)) return rc;
    const int nb = eikonal_batch(c, ny, nx, nsrc);
    EikConsts kc;
    if (int rc = eikonal_consts(c, dx, dy, &kc, nullptr)) return rc;
    const int *sx, *sz;
    const float *va, *vm;
    float *work;
    if (int rc = eikonal_setup(c, d_vel, ny, nx, nsrc, src_x, src_z, nb, &sx, &sz, &va, &vm, &work)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        if (int rc = eikonal_events(c)) return rc;
    const size_t ncell = (size_t)nx * ny;
    // batches of nb sources reuse one workspace; stream order serialises them
    for (int b0 = 0; b0 < nsrc; b0 += nb) {
        const int n = std::min(nb, nsrc - b0);
        eikonal_launch(c, va, vm, ny, nx, kc, n, sx + b0, sz + b0, work, d_tt + (size_t)b0 * ncell,
                       timed && b0 == 0 ? c->kir_ev[0] : nullptr, timed && b0 + n == nsrc ? c->kir_ev[1] : nullptr);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (timed)
        if (int rc = eikonal_time(c, true)) return rc;
    return 0;
}

extern "C" int psdm_rtm_eikonal_host(psdm_rtm_ctx *c, const float *h_vel, int ny, int nx, float dx, float dy,
                                     int nsrc, const int *src_x, const int *src_z, float *h_tt) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!h_vel || !h_tt) return fail(c, PSDM_RTM_EINVAL, "eikonal_host: null argument");
    if (int rc = check_eikonal(c, ny, nx, dx, dy, nsrc, src_x, src_z)) return rc;
    if (int rc = set_device(c)) return rc;
    // the device holds one batch of travel times at a time; each is copied out before the next
    const int nb = eikonal_batch(c, ny, nx, nsrc);
    void *dv, *dt;
    const size_t n = (size_t)nx * ny;
    if (int rc = get_buf(c, "eikh_vel", n * sizeof(float), &dv)) return rc;
    if (int rc = get_buf(c, "eikh_tt", n * (size_t)nb * sizeof(float), &dt)) return rc;
    HIP_TRY(c, hipMemcpyAsync(dv, h_vel, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    EikConsts kc;
    if (int rc = eikonal_consts(c, dx, dy, &kc, nullptr)) return rc;
    const int *sx, *sz;
    const float *va, *vm;
    float *work;
    if (int rc = eikonal_setup(c, (const float *)dv, ny, nx, nsrc, src_x, src_z, nb, &sx, &sz, &va, &vm, &work))
        return rc;
    const bool timed = c->timing != 0;
    if (timed)
        if (int rc = eikonal_events(c)) return rc;
    for (int b0 = 0; b0 < nsrc; b0 += nb) {
        const int m = std::min(nb, nsrc - b0);
        eikonal_launch(c, va, vm, ny, nx, kc, m, sx + b0, sz + b0, work, (float *)dt, timed ? c->kir_ev[0] : nullptr,
                       timed ? c->kir_ev[1] : nullptr);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(h_tt + (size_t)b0 * n, dt, n * (size_t)m * sizeof(float), hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (timed)
            if (int rc = eikonal_time(c, b0 == 0)) return rc;
    }
    return 0;
}

extern "C" int psdm_rtm_set_eikonal_workspace(psdm_rtm_ctx *c, size_t max_bytes) {
    if (!c) return PSDM_RTM_EINVAL;
    if (max_bytes == 0) return fail(c, PSDM_RTM_EINVAL, "eikonal workspace limit must be > 0");
    c->eik_work_cap = max_bytes;
    return 0;
}

extern "C" int psdm_rtm_eikonal_division(psdm_rtm_ctx *c, float dx, float dy, int *fast_mask) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!fast_mask || !(dx > 0.f) || !(dy > 0.f)) return fail(c, PSDM_RTM_EINVAL, "eikonal_division: bad argument");
    if (int rc = set_device(c)) return rc;
    EikConsts kc;
    return eikonal_consts(c, dx, dy, &kc, fast_mask);
}

extern "C" int psdm_rtm_eikonal_batch(psdm_rtm_ctx *c, int ny, int nx, int nsrc, int *nb) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!nb || ny < 1 || nx < 1 || nsrc < 1) return fail(c, PSDM_RTM_EINVAL, "eikonal_batch: bad argument");
    *nb = eikonal_batch(c, ny, nx, nsrc);
    return 0;
}

extern "C" int psdm_rtm_kirchhoff_image(psdm_rtm_ctx *c, const float *d_tts, const float *d_ttr, const float *d_data,
                                        int ld, int nrec, int nt_prop, int ny, int nx, float dd, float dt, float vp0,
                                        float theta_max, float theta_taper, float wmax, float *d_img) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_tts || !d_ttr || !d_data || !d_img) return fail(c, PSDM_RTM_EINVAL, "kirchhoff_image: null argument");
    if (ny < 1 || nx < 1 || nrec < 1 || nt_prop < 1 || ld < nt_prop)
        return fail(c, PSDM_RTM_EINVAL, "kirchhoff_image: need ny, nx, nrec, nt_prop >= 1 and ld >= nt_prop");
    if ((long long)nx * ny > 2147483647LL - 256) return fail(c, PSDM_RTM_EINVAL, "kirchhoff_image: image too large");
    if (!(dt > 0.f)) return fail(c, PSDM_RTM_EINVAL, "kirchhoff_image: dt must be > 0");
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (int q = 2; q < 4; ++q)
            if (!c->kir_ev[q]) HIP_TRY(c, hipEventCreate(&c->kir_ev[q]));
    ImgConsts kc;
    {
        const float ys[3] = {dt, 2.0f * dt, dt};
        DivSet ds;
        int mask;
        if (int rc = fast_divisors(c, ys, &ds, &mask)) return rc;
        kc.zh_t = ds.zh[0];
        kc.zl_t = ds.zl[0];
        kc.zh_2t = ds.zh[1];
        kc.zl_2t = ds.zl[1];
        kc.fast = mask & 3;
        kc.cos_in = (float)(std::cos((double)theta_max * 3.14159265358979323846 / 180.0) + 1e-4);
    }
    const int np = nx * ny;
    hipExtLaunchKernelGGL(kirchhoff_image_kernel, dim3((np + 255) / 256), dim3(256), 0, c->stream,
                          timed ? c->kir_ev[2] : nullptr, timed ? c->kir_ev[3] : nullptr, 0, d_tts, d_ttr, d_data, ld,
                          nrec, nt_prop, ny, nx, dd, dt, vp0, theta_max, theta_taper, wmax, kc, d_img);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (timed) {
        float ms;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->kir_ev[2], c->kir_ev[3]));
        c->kir_ms[1] += ms;
        c->kir_n[1]++;
    }
    return 0;
}

extern "C" int psdm_rtm_kirchhoff_timing(psdm_rtm_ctx *c, double *eikonal_ms, double *image_ms, long long *n_eikonal,
                                         long long *n_image) {
    if (!c) return PSDM_RTM_EINVAL;
    if (eikonal_ms) *eikonal_ms = c->kir_n[0] ? c->kir_ms[0] / (double)c->kir_n[0] : 0.0;
    if (image_ms) *image_ms = c->kir_n[1] ? c->kir_ms[1] / (double)c->kir_n[1] : 0.0;
    if (n_eikonal) *n_eikonal = c->kir_n[0];
    if (n_image) *n_image = c->kir_n[1];
    return 0;
}
