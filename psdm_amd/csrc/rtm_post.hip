// Depth-to-two-way-time conversion of the final image on the device: psdm2twt
// (auxiliar_PSDM.f90:23-91) without its file writers.
//
//   twt_times_kernel   one thread per model column (depth chunks staged in LDS so the loads
//                      and stores coalesce): the two-way time at each depth
//                      sample accumulated in FP32 in depth order, t(1) = 0,
//                      t(k) = t(k-1) + 2*dd/v(k) where |v(k)| > 0 (:49-60), the column's
//                      max |t| (for nt = 1 + ceiling(max/dt), :65-67) and whether t is
//                      non-decreasing;
//   twt_interp_kernel  one thread per (time sample, column): linear_interpolation_1D
//                      (linear_interpolation.f90:72-103) of the velocity and of the image
//                      at t = (j-1)*dt -- constant beyond the ends, inside the first
//                      interval containing t whose two ordinates are non-zero, 0 if none --
//                      with the reference's FP32 operations.  Non-decreasing columns find
//                      that interval by bisection, the others scan as the reference does.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "rtm_ctx.hpp"

using psdm::host::fail;
using psdm::host::get_buf;
using psdm::host::set_device;

namespace psdm {
namespace {

// 64 columns per workgroup; depth runs in chunks of 128 samples staged in LDS: all 256
// threads load a chunk (consecutive lanes on consecutive depths of a column) and turn
// each velocity into its increment 2*dd/v (0 where v = 0), then the first wavefront adds
// the increments of its column in depth order (the only serial part), and all threads
// store the times back.
constexpr int kTC = 64, kTK = 128;
__global__ __launch_bounds__(256) void twt_times_kernel(const float *__restrict__ vel, int nz, int nx, float dd,
                                                        float *__restrict__ t, int *__restrict__ mono,
                                                        unsigned *__restrict__ tmax_bits) {
    __shared__ float tile[kTC][kTK + 1];
    const int c0 = blockIdx.x * kTC, tid = threadIdx.x;
    const float two_dd = 2.0f * dd;
    float acc = 0.f, m = 0.f;
    int up = 1;
    for (int k0 = 0; k0 < nz; k0 += kTK) {
        for (int e = tid; e < kTC * kTK; e += 256) {
            const int col = e / kTK, kk = e % kTK, k = k0 + kk;
            float inc = 0.f;
            if (c0 + col < nx && k < nz && k > 0) {  // t(1) = 0: no increment at the first sample
                const float vk = vel[(size_t)(c0 + col) * nz + k];
                if (fabsf(vk) > 0.f) inc = two_dd / vk;
            }
            tile[col][kk] = inc;
        }
        __syncthreads();
        if (tid < kTC) {
#pragma unroll 8
            for (int kk = 0; kk < kTK; ++kk) {
                const float nv = acc + tile[tid][kk];  // where v = 0 the reference keeps t(k-1): acc + 0 is acc
                if (!(nv >= acc)) up = 0;
                acc = nv;
                m = fmaxf(m, fabsf(acc));
                tile[tid][kk] = acc;
            }
        }
        __syncthreads();
        for (int e = tid; e < kTC * kTK; e += 256) {
            const int col = e / kTK, kk = e % kTK, k = k0 + kk;
            if (c0 + col < nx && k < nz) t[(size_t)(c0 + col) * nz + k] = tile[col][kk];
        }
        __syncthreads();
    }
    if (tid < kTC && c0 + tid < nx) {
        mono[c0 + tid] = up;
        atomicMax(tmax_bits, __float_as_uint(m));  // non-negative floats order as their bit patterns
    }
}

__device__ float lin_interp(const float *__restrict__ x1, const float *__restrict__ y1, int n1, float a3, bool up) {
    if (a3 <= x1[0]) return y1[0];
    if (a3 >= x1[n1 - 1]) return y1[n1 - 1];
    int i = 0;
    if (up) {  // first interval whose right end reaches a3; those before it cannot contain a3
        int lo = 1, hi = n1 - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (x1[mid] >= a3)
                hi = mid;
            else
                lo = mid + 1;
        }
        i = lo - 1;
    }
    for (; i < n1 - 1; ++i) {
        const float a1 = x1[i + 1], a2 = x1[i], b1 = y1[i + 1], b2 = y1[i];
        if (a3 >= a2 && a3 <= a1 && fabsf(b1) > 0.f && fabsf(b2) > 0.f) {
            const float m = (b1 - b2) / (a1 - a2);
            return b2 + (a3 - a2) * m;
        }
        if (up && a2 > a3) break;  // every later interval starts beyond a3
    }
    return 0.f;
}

__global__ __launch_bounds__(256) void twt_interp_kernel(const float *__restrict__ t, const int *__restrict__ mono,
                                                         const float *__restrict__ vel, const float *__restrict__ psdm,
                                                         int nz, int nt, float dt, float *__restrict__ vel_twt,
                                                         float *__restrict__ psdm_twt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= nt) return;
    const float a3 = (float)j * dt;
    const size_t c = (size_t)i * nz;
    const bool up = mono[i] != 0;
    vel_twt[(size_t)i * nt + j] = lin_interp(t + c, vel + c, nz, a3, up);
    psdm_twt[(size_t)i * nt + j] = lin_interp(t + c, psdm + c, nz, a3, up);
}

__global__ void twt_clear_kernel(unsigned *p) {
    if (threadIdx.x == 0) *p = 0u;
}

// Scratch of psdm_rtm_psdm2twt_noalloc, one copy per GPU held until the process ends: the
// two-way times (nz*nx floats), then at the next 256-byte boundary the per-column
// monotonicity flags (nx ints) and the max |t| bits (one word).  Host code reaches it only
// through the per-device address from hipGetSymbolAddress (scratch_base).
constexpr size_t kTwtScratchBytes = size_t(32) << 20;
__device__ __align__(256) unsigned char scratch[kTwtScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (the context's device is current): looked up once per
// device and kept.
int scratch_base(psdm_rtm_ctx *c, unsigned char **out) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, PSDM_RTM_EINVAL, "psdm2twt_noalloc: device id");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return psdm::host::hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[c->device] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[c->device];
    return 0;
}

size_t twt_scratch_need(int nx, int nz, size_t *mono_off) {
    *mono_off = ((size_t)nz * nx * sizeof(float) + 255) & ~(size_t)255;
    return *mono_off + ((size_t)nx + 1) * sizeof(int);
}

int check_twt_args(psdm_rtm_ctx *c, int nx, int nz, float dd, float dt, const float *d_vel, const float *d_psdm,
                   const int *nt) {
    if (!d_vel || !d_psdm || !nt) return fail(c, PSDM_RTM_EINVAL, "psdm2twt: null argument");
    if (nx < 1 || nx > 65535 || nz < 1) return fail(c, PSDM_RTM_EINVAL, "psdm2twt: need 1 <= nx <= 65535, nz >= 1");
    if (!(dd > 0.f) || !(dt > 0.f)) return fail(c, PSDM_RTM_EINVAL, "psdm2twt: dd and dt must be > 0");
    return 0;
}

// The conversion with its workspace at d_t (nz*nx floats) and d_mono (nx ints, then the max
// bits): the times, nt, and, unless d_vel_twt or d_psdm_twt is null, the interpolation.
// `timed`: device times of both kernels into the context.  Synchronous.
int twt_run(psdm_rtm_ctx *c, int nx, int nz, float dd, float dt, const float *d_vel, const float *d_psdm, int nt_cap,
            float *d_vel_twt, float *d_psdm_twt, int *nt, float *d_t, int *d_mono, bool timed) {
    unsigned *d_max = (unsigned *)(d_mono + nx);
    hipLaunchKernelGGL(twt_clear_kernel, dim3(1), dim3(64), 0, c->stream, d_max);
    hipExtLaunchKernelGGL(twt_times_kernel, dim3((nx + kTC - 1) / kTC), dim3(256), 0, c->stream,
                          timed ? c->twt_ev[0] : nullptr, timed ? c->twt_ev[1] : nullptr, 0, d_vel, nz, nx, dd, d_t,
                          d_mono, d_max);
    HIP_TRY(c, hipGetLastError());
    unsigned bits = 0;
    HIP_TRY(c, hipMemcpyAsync(&bits, d_max, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float tmax;
    std::memcpy(&tmax, &bits, sizeof(float));
    const double nd = 1.0 + std::ceil((double)(tmax / dt));  // nt = 1 + ceiling(time_max/dt), quotient in FP32
    if (!(nd < 2147483647.0)) return fail(c, PSDM_RTM_EINVAL, "psdm2twt: time axis too long (is dt or v wrong?)");
    *nt = (int)nd;
    if (!d_vel_twt || !d_psdm_twt) return 0;  // length query
    if (*nt > nt_cap) return fail(c, PSDM_RTM_EINVAL, "psdm2twt: outputs need nt rows (nt_cap too small)");
    hipExtLaunchKernelGGL(twt_interp_kernel, dim3((*nt + 255) / 256, nx), dim3(256), 0, c->stream,
                          timed ? c->twt_ev[2] : nullptr, timed ? c->twt_ev[3] : nullptr, 0, (const float *)d_t,
                          (const int *)d_mono, d_vel, d_psdm, nz, *nt, dt, d_vel_twt, d_psdm_twt);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (timed) {
        float a, b;
        HIP_TRY(c, hipEventElapsedTime(&a, c->twt_ev[0], c->twt_ev[1]));
        HIP_TRY(c, hipEventElapsedTime(&b, c->twt_ev[2], c->twt_ev[3]));
        c->twt_ms[0] += a;
        c->twt_ms[1] += b;
        c->twt_n++;
    }
    return 0;
}

}  // namespace
}  // namespace psdm

using namespace psdm;

extern "C" int psdm_rtm_psdm2twt(psdm_rtm_ctx *c, int nx, int nz, float dd, float dt, const float *d_vel,
                                 const float *d_psdm, int nt_cap, float *d_vel_twt, float *d_psdm_twt, int *nt) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = check_twt_args(c, nx, nz, dd, dt, d_vel, d_psdm, nt)) return rc;
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (auto &e : c->twt_ev)
            if (!e) HIP_TRY(c, hipEventCreate(&e));
    size_t mono_off;
    void *ws;
    if (int rc = get_buf(c, "twt_ws", twt_scratch_need(nx, nz, &mono_off), &ws)) return rc;
    return twt_run(c, nx, nz, dd, dt, d_vel, d_psdm, nt_cap, d_vel_twt, d_psdm_twt, nt, (float *)ws,
                   (int *)((unsigned char *)ws + mono_off), timed);
}

extern "C" int psdm_rtm_psdm2twt_noalloc(psdm_rtm_ctx *c, int nx, int nz, float dd, float dt, const float *d_vel,
                                         const float *d_psdm, int nt_cap, float *d_vel_twt, float *d_psdm_twt,
                                         int *nt) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = check_twt_args(c, nx, nz, dd, dt, d_vel, d_psdm, nt)) return rc;
    size_t mono_off;
    if (twt_scratch_need(nx, nz, &mono_off) > kTwtScratchBytes)
        return fail(c, PSDM_RTM_EINVAL, "psdm2twt_noalloc: nz*nx + nx exceeds the scratch (32 MiB: about 8.3 M cells)");
    if (int rc = set_device(c)) return rc;
    unsigned char *base;
    if (int rc = scratch_base(c, &base)) return rc;
    return twt_run(c, nx, nz, dd, dt, d_vel, d_psdm, nt_cap, d_vel_twt, d_psdm_twt, nt, (float *)base,
                   (int *)(base + mono_off), false);
}

extern "C" int psdm_rtm_psdm2twt_timing(psdm_rtm_ctx *c, double *times_ms, double *interp_ms, long long *n) {
    if (!c) return PSDM_RTM_EINVAL;
    if (times_ms) *times_ms = c->twt_n ? c->twt_ms[0] / (double)c->twt_n : 0.0;
    if (interp_ms) *interp_ms = c->twt_n ? c->twt_ms[1] / (double)c->twt_n : 0.0;
    if (n) *n = c->twt_n;
    return 0;
}
