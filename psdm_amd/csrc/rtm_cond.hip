// Gather conditioning before migration on the device: the driver's per-shot steps
// between reading a gather and migrating it (calculate_PSDM_RTM.f90:299-358).
//
//   filter     freq_filter (auxiliar.f90:498-531): each trace is zero-padded to
//              N = 24*nt samples with the trace at the end, transformed (FFTPACK cfft1f,
//              normalised by 1/N), multiplied by the Butterworth transfer function H of
//              filters_function1D (filters.f90:126-232), transformed back (cfft1b) and the
//              real part of its last nt samples kept.  That is a linear, shift-invariant
//              operator on the trace's own samples: y[n] = sum_m g[n-m] x[m] with
//              g = Re(IDFT_N(H)) at lags -(nt-1)..nt-1.  The taps are computed once per
//              (nt, dt, filter) in FP64 (cond_taps_kernel) from H evaluated at the
//              reference's own FP32-accumulated frequencies, and every gather is then one
//              Toeplitz product over all its traces on the matrix cores
//              (cond_toeplitz_kernel, FP32 MFMA).
//   resample   interpolation_akima (interpolation_akima.f90:1-31) over INTRPL
//              (akima433.f90:6-170), one thread per output sample, FP32 operations in the
//              reference's order (bit for bit); fused with
//   cut/taper  cut_sf zeroing (:325-329), taper_pad_data / taper_pad_trace with the
//              cut_multiple zeroing (:331-346, filters.f90:1-48); the taper weights are
//              computed on the host with libm's cosf, as the reference's cos.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "rtm_ctx.hpp"

using psdm::host::fail;
using psdm::host::get_buf;
using psdm::host::set_device;

namespace psdm {
namespace {

constexpr int kZeroPad = 24;  // auxiliar.f90:507

// g[d + nt - 1] = (1/N) [Re H_0 + 2 sum_{k=1}^{N/2-1} Re(H_k e^{2 pi i k d / N}) + Re H_{N/2} (-1)^d]
// for d = -(nt-1)..nt-1: the real part of the inverse DFT of the full H (H_{N-k} =
// conj(H_k), filters.f90:227-229), one workgroup per lag, FP64.
__global__ __launch_bounds__(256) void cond_taps_kernel(const double *reH, const double *imH, int N, int nt, float *g) {
    const int d = (int)blockIdx.x - (nt - 1);
    const long long ad = d < 0 ? -(long long)d : (long long)d;
    const double sg = d < 0 ? -1.0 : 1.0;  // sin(-x) = -sin(x)
    double acc = 0.0;
    for (int k = 1 + (int)threadIdx.x; k < N / 2; k += (int)blockDim.x) {
        const long long r = ((long long)k * ad) % N;
        double s, c;
        sincospi(2.0 * (double)r / (double)N, &s, &c);
        acc += reH[k] * c - imH[k] * (sg * s);
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double nyq = reH[N / 2] * ((ad & 1) ? -1.0 : 1.0);
        g[blockIdx.x] = (float)((reH[0] + 2.0 * red[0] + nyq) / (double)N);
    }
}

// y[r*nt + n] = sum_{m=0}^{nt-1} g[n - m + nt - 1] x[r*nt + m]: the gather times the
// Toeplitz matrix T[n][m] = g[n - m] on the matrix cores.  A workgroup computes a
// 64-sample x 64-trace tile, each wavefront 32 x 32 with v_mfma_f32_32x32x2f32
// (A = T rows n, columns m..m+1: lane l holds T[i = l % 32][k = l / 32]; B = x rows
// m..m+1: lane l holds x[k = l / 32][j = l % 32]; the 16 accumulators of lane l are
// C[8*(v/4) + 4*(l/32) + v%4][l % 32]).  m runs in ascending chunks of 64 staged in LDS
// with the 127 taps they need; the next chunk's global loads are issued into registers
// before the current chunk's MFMAs, so their latency overlaps the matrix work.
constexpr int kTN = 64, kTR = 64, kTM = 64, kTG = kTN + kTM - 1;
typedef float v16f __attribute__((ext_vector_type(16)));
// Split over m: workgroup part kp of a tile sums the chunks [kp*cpp, (kp+1)*cpp) of 64 and
// writes its partial product to part[kp] (ksplit > 1; cond_reduce_kernel adds the parts in
// order), so a gather of few tiles still spreads over all CUs.
__global__ __launch_bounds__(256) void cond_toeplitz_kernel(const float *__restrict__ x, int nt, int nrec,
                                                            const float *__restrict__ g, float *__restrict__ y,
                                                            int n_ntile, int n_tiles, int per_xcd, int ksplit,
                                                            int cpp, float *__restrict__ part) {
    __shared__ float xs[kTM][kTR + 1];
    __shared__ float gs[kTG];
    // XCD-aware order: the 8 XCDs take workgroups round-robin, so XCD b % 8 gets the
    // contiguous tile range [(b % 8) * per_xcd, ...), tiles trace-block-major: all sample
    // tiles of a trace block run on one XCD and read its traces through one L2
    const int tt = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tt >= n_tiles * ksplit) return;
    const int t = tt / ksplit, kp = tt - t * ksplit;  // the parts of a tile are neighbours
    const int n0 = (t % n_ntile) * kTN, r0 = (t / n_ntile) * kTR;
    const int mb = kp * cpp * kTM, me = min(nt, mb + cpp * kTM);
    float *out = ksplit > 1 ? part + (size_t)kp * nrec * nt : y;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wn = 32 * (w & 1), wr = 32 * (w >> 1);  // this wavefront's 32 x 32 sub-tile
    const int li = lane & 31, lk = lane >> 5;
    constexpr int kPer = kTM * kTR / 256;  // x values staged per thread and chunk
    float xr[kPer], gr = 0.f;
    auto fetch = [&](int m0) {  // element e = tid + 256 q: row m0 + e % kTM, trace r0 + e / kTM
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q, m = m0 + e % kTM, r = r0 + e / kTM;
            xr[q] = (m < nt && r < nrec) ? x[(size_t)r * nt + m] : 0.f;
        }
        if (tid < kTG) {
            const int d = n0 - m0 - (kTM - 1) + tid;  // lag n - m of this entry
            gr = (d > -nt && d < nt) ? g[d + nt - 1] : 0.f;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q;
            xs[e % kTM][e / kTM] = xr[q];
        }
        if (tid < kTG) gs[tid] = gr;
    };
    // four independent accumulators (m steps 8q, 8q+2, 8q+4, 8q+6): a 32x32x2 MFMA waits
    // for the previous one on the same accumulator, and with ~1 wavefront per SIMD (304
    // workgroups for a 480-trace gather) one chain would leave the matrix core idle
    v16f acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[q][v] = 0.f;
    fetch(mb);
    stage();
    __syncthreads();
    for (int m0 = mb; m0 < me; m0 += kTM) {
        const bool more = m0 + kTM < me;
        if (more) fetch(m0 + kTM);
#pragma unroll 2
        for (int m8 = 0; m8 < kTM; m8 += 8)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int mm = m8 + 2 * q;
                const float a = gs[wn + li - (mm + lk) + kTM - 1];  // T[n0 + wn + li][m0 + mm + lk]
                const float b = xs[mm + lk][wr + li];               // x[m0 + mm + lk][r0 + wr + li]
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[q], 0, 0, 0);
            }
        __syncthreads();
        if (more) {
            stage();
            __syncthreads();
        }
    }
    const int r = r0 + wr + li;
    if (r < nrec)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int n = n0 + wn + 8 * (v >> 2) + 4 * lk + (v & 3);
            if (n < nt) out[(size_t)r * nt + n] = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
        }
}

__global__ void cond_reduce_kernel(const float *__restrict__ part, int ksplit, size_t n, float *__restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float acc = part[e];
    for (int k = 1; k < ksplit; ++k) acc = acc + part[(size_t)k * n + e];
    y[e] = acc;
}

// INTRPL (akima433.f90:6-170) at one abscissa uk for the samples y(1..L) at
// x(i) = (i-1)*dx: the interval search of :48-65 and the coefficients of :66-131, in the
// reference's FP32 operation order.  A2 and A4 keep their initial 0 where the reference
// leaves them unassigned (L = 2).
__device__ float akima_at(const float *__restrict__ y, int L, float dx, float uk) {
    auto X = [&](int i) { return (float)(i - 1) * dx; };
    auto Y = [&](int i) { return y[i - 1]; };
    const int LM1 = L - 1, LM2 = L - 2, LP1 = L + 1;
    int I;
    if (LM2 == 0) {
        I = 2;
    } else if (uk >= X(L)) {
        I = LP1;
    } else if (uk < X(1)) {
        I = 1;
    } else {
        int imn = 2, imx = L;
        do {
            I = (imn + imx) / 2;
            if (uk >= X(I))
                imn = I + 1;
            else
                imx = I;
        } while (imx > imn);
        I = imx;
    }
    int J = I;
    if (J == 1) J = 2;
    if (J == LP1) J = L;
    float X3 = X(J - 1), Y3 = Y(J - 1);
    const float X4 = X(J), Y4 = Y(J);
    float A3 = X4 - X3;
    float M3 = (Y4 - Y3) / A3;
    float X2 = 0.f, Y2 = 0.f, A2 = 0.f, M2, X5 = 0.f, Y5 = 0.f, A4 = 0.f, M4, M1, M5;
    if (LM2 == 0) {
        M2 = M3;
        M4 = M3;
    } else {
        bool to42 = false;
        if (J != 2) {
            X2 = X(J - 2);
            Y2 = Y(J - 2);
            A2 = X3 - X2;
            M2 = (Y3 - Y2) / A2;
            to42 = J == L;
        }
        if (!to42) {
            X5 = X(J + 1);
            Y5 = Y(J + 1);
            A4 = X5 - X4;
            M4 = (Y5 - Y4) / A4;
            if (J == 2) M2 = M3 + M3 - M4;
        } else {
            M4 = M3 + M3 - M2;
        }
    }
    if (J <= 3) {
        M1 = M2 + M2 - M3;
    } else {
        const float A1 = X2 - X(J - 3);
        M1 = (Y2 - Y(J - 3)) / A1;
    }
    if (J >= LM1) {
        M5 = M4 + M4 - M3;
    } else {
        const float A5 = X(J + 2) - X5;
        M5 = (Y(J + 2) - Y5) / A5;
    }
    float T3 = 0.f, T4;
    if (I != LP1) {
        float W2 = fabsf(M4 - M3), W3 = fabsf(M2 - M1), SW = W2 + W3;
        if (SW == 0.f) {
            W2 = 0.5f;
            W3 = 0.5f;
            SW = 1.0f;
        }
        T3 = (W2 * M2 + W3 * M3) / SW;
    }
    if (I == 1) {
        T4 = T3;
        const float SA = A3 + A4;
        T3 = 0.5f * (M1 + M2 - A4 * (A3 - A4) * (M3 - M4) / (SA * SA));
        X3 = X3 - A4;
        Y3 = Y3 - M2 * A4;
        A3 = A4;
        M3 = M2;
    } else {
        float W3 = fabsf(M5 - M4), W4 = fabsf(M3 - M2), SW = W3 + W4;
        if (SW == 0.f) {
            W3 = 0.5f;
            W4 = 0.5f;
            SW = 1.0f;
        }
        T4 = (W3 * M3 + W4 * M4) / SW;
        if (I == LP1) {
            T3 = T4;
            const float SA = A2 + A3;
            T4 = 0.5f * (M4 + M5 - A2 * (A2 - A3) * (M2 - M3) / (SA * SA));
            X3 = X4;
            Y3 = Y4;
            A3 = A2;
            M3 = M4;
        }
    }
    const float Q2 = (2.0f * (M3 - T3) + M3 - T4) / A3;
    const float Q3 = (-M3 - M3 + T3 + T4) / (A3 * A3);
    const float DX = uk - X3;
    return Y3 + DX * (T3 + DX * (Q2 + DX * Q3));
}

// out[r*nt_taper + k], k < nt_taper: resampled (or copied) sample k+1 of trace r, then
// the cut_sf zeroing and the taper / cut_multiple of the driver.  mode 0: no taper,
// 1: taper ending at nt_prop2, 2: cut at nt_cut(r) with the taper ending there.
__global__ __launch_bounds__(256) void cond_resample_kernel(const float *__restrict__ src, int L, float dtd,
                                                            int resample, int nt_prop2, float dt, int nt_taper,
                                                            int mode, int N_taper, const float *__restrict__ w,
                                                            const int *__restrict__ nt_sf,
                                                            const int *__restrict__ nt_cut, float *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (k >= nt_taper) return;
    const float *tr = src + (size_t)r * L;
    float v = 0.f;
    if (k < nt_prop2) v = resample ? akima_at(tr, L, dtd, (float)k * dt) : tr[k];
    const int s = k + 1;  // 1-based sample
    if (nt_sf && s <= nt_sf[r]) v = 0.f;
    if (mode == 1) {
        const int i0 = std::max(1, nt_prop2 - N_taper + 1), q = s - i0;
        if (q >= 0 && q < N_taper && s <= nt_prop2) v = v * w[q];
    } else if (mode == 2) {
        const int c = nt_cut[r];
        if (s >= c) {
            v = 0.f;
        } else {
            const int i0 = std::max(1, c - N_taper + 1), q = s - i0;
            if (q >= 0 && q < N_taper) v = v * w[q];
        }
    }
    out[(size_t)r * nt_taper + k] = v;
}

unsigned bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}

// H(j), j = 1..ts/2+1 of filters_function1D (filters.f90:71-232) with n = 6, parity 0:
// frequencies accumulated in FP32 as the reference (f = f + fe/ts), sections in FP64.
void butterworth(int ts, float fe, int typef, float f1, float fc, std::vector<double> &re, std::vector<double> &im) {
    const int n = 6, nh = ts / 2 + 1;
    const float pi = 3.14159265f;
    const double fs = 40.0;
    const float f0f = std::sqrt(f1 * fc);
    const double f0 = f0f, Q = (double)(f0f / (fc - f1));
    re.assign(nh, 0.0);
    im.assign(nh, 0.0);
    const float step = fe / (float)ts;
    float f = 0.f;
    double cs[3];
    for (int k = 1; k <= n / 2; ++k) cs[k - 1] = std::cos((double)((float)(2 * k + n - 1) * pi / (float)(2 * n)));
    for (int j = 0; j < nh; ++j) {
        const double fd = f;
        std::complex<double> B(1.0, 0.0);
        if (j > 0 || typef == 1) {
            std::complex<double> z;
            const std::complex<double> I1(0.0, 1.0);
            if (typef == 1)
                z = I1 * fd / (double)fc;
            else if (typef == 2)
                z = fs / (I1 * fd);
            else
                z = Q * ((I1 * fd) / f0 + f0 / (I1 * fd));
            for (int k = 0; k < n / 2; ++k) B *= z * z - 2.0 * z * cs[k] + 1.0;
        }
        const std::complex<double> H = 1.0 / B;
        re[j] = H.real();
        im[j] = H.imag();
        f = f + step;
    }
    re[0] = typef == 1 ? 1.0 : 0.0;  // :221-225
    im[0] = 0.0;
}

}  // namespace
}  // namespace psdm

using namespace psdm;

extern "C" int psdm_rtm_condition(psdm_rtm_ctx *c, const psdm_rtm_cond *q, int NumRec, const float *d_field,
                                  float *d_target) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!q || !d_field || !d_target) return fail(c, PSDM_RTM_EINVAL, "condition: null argument");
    const bool filt = q->f_init > 0.f && q->f_final > 0.f;
    const bool resample = q->dt_data != q->dt_prop;
    if (NumRec < 1 || NumRec > 65535 || q->nt_data < 1 || q->nt_prop2 < 1 || q->nt_taper < q->nt_prop2)
        return fail(c, PSDM_RTM_EINVAL,
                    "condition: need 1 <= NumRec <= 65535, nt_data, nt_prop2 >= 1 and nt_taper >= nt_prop2");
    if (!(q->dt_data > 0.f) || !(q->dt_prop > 0.f)) return fail(c, PSDM_RTM_EINVAL, "condition: dt must be > 0");
    if (resample && q->nt_data < 2) return fail(c, PSDM_RTM_EINVAL, "condition: resampling needs nt_data >= 2");
    if (!resample && q->nt_prop2 > q->nt_data)
        return fail(c, PSDM_RTM_EINVAL, "condition: nt_prop2 > nt_data without resampling");
    if (filt && (q->typef < 1 || q->typef > 3)) return fail(c, PSDM_RTM_EINVAL, "condition: typef must be 1, 2 or 3");
    if (filt && (long long)kZeroPad * q->nt_data > (1LL << 30))
        return fail(c, PSDM_RTM_EINVAL, "condition: nt_data too large for the filter");
    if ((long long)q->nt_taper * NumRec > (1LL << 31) || (long long)q->nt_data * NumRec > (1LL << 31))
        return fail(c, PSDM_RTM_EINVAL, "condition: gather too large");
    const int mode = q->nt_cut ? 2 : (q->taper_option != 0 ? 1 : 0);
    if (mode != 0 && q->N_taper < 1) return fail(c, PSDM_RTM_EINVAL, "condition: N_taper must be >= 1");
    if (q->nt_cut)
        for (int i = 0; i < NumRec; ++i)
            if (q->nt_cut[i] < 1 || q->nt_cut[i] > q->nt_taper)
                return fail(c, PSDM_RTM_EINVAL, "condition: nt_cut out of 1..nt_taper");
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (auto &e : c->cond_ev)
            if (!e) HIP_TRY(c, hipEventCreate(&e));
    const int nt = q->nt_data;
    const float *src = d_field;
    if (filt) {
        const float fe = 1.0f / q->dt_data;  // auxiliar.f90:509
        const std::vector<unsigned> key = {(unsigned)nt, bits(fe), (unsigned)q->typef, bits(q->f_init),
                                           bits(q->f_final)};
        void *vt;
        if (int rc = get_buf(c, "cond_taps", (size_t)(2 * nt - 1) * sizeof(float), &vt)) return rc;
        if (c->cond_key != key) {
            const int N = kZeroPad * nt;
            std::vector<double> re, im;
            butterworth(N, fe, q->typef, q->f_init, q->f_final, re, im);
            void *vh;
            if (int rc = get_buf(c, "cond_h", 2 * re.size() * sizeof(double), &vh)) return rc;
            double *dre = (double *)vh, *dim = dre + re.size();
            HIP_TRY(c, hipMemcpyAsync(dre, re.data(), re.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(dim, im.data(), im.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL(cond_taps_kernel, dim3(2 * nt - 1), dim3(256), 0, c->stream, dre, dim, N, nt, (float *)vt);
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, hipStreamSynchronize(c->stream));  // re / im are host locals
            c->cond_key = key;
        }
        void *vf;
        if (int rc = get_buf(c, "cond_filtered", (size_t)nt * NumRec * sizeof(float), &vf)) return rc;
        const int n_ntile = (nt + kTN - 1) / kTN, n_tiles = n_ntile * ((NumRec + kTR - 1) / kTR);
        // split the m sums so that ~1024 workgroups run (4 per CU), each over >= 4 chunks
        const int nchunk = (nt + kTM - 1) / kTM;
        int ksplit = std::max(1, std::min({8, (1024 + n_tiles - 1) / n_tiles, nchunk / 4}));
        const int cpp = (nchunk + ksplit - 1) / ksplit;
        ksplit = (nchunk + cpp - 1) / cpp;
        const int per_xcd = (n_tiles * ksplit + 7) / 8;
        void *vp = nullptr;
        const size_t n = (size_t)nt * NumRec;
        if (ksplit > 1)
            if (int rc = get_buf(c, "cond_parts", (size_t)ksplit * n * sizeof(float), &vp)) return rc;
        hipExtLaunchKernelGGL(cond_toeplitz_kernel, dim3(8 * per_xcd), dim3(256), 0, c->stream,
                              timed ? c->cond_ev[0] : nullptr, (timed && ksplit == 1) ? c->cond_ev[1] : nullptr, 0,
                              d_field, nt, NumRec, (const float *)vt, (float *)vf, n_ntile, n_tiles, per_xcd, ksplit, cpp,
                              (float *)vp);
        HIP_TRY(c, hipGetLastError());
        if (ksplit > 1) {
            hipExtLaunchKernelGGL(cond_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, nullptr,
                                  timed ? c->cond_ev[1] : nullptr, 0, (const float *)vp, ksplit, n, (float *)vf);
            HIP_TRY(c, hipGetLastError());
        }
        src = (const float *)vf;
    }
    // taper weights w(n) = 0.5*(1 + cos(pi*n/N_taper)), pi = 4 atan(1) (filters.f90:32-40)
    std::vector<float> w;
    if (mode != 0) {
        const float pi = 4.0f * std::atan(1.0f);
        w.resize(q->N_taper);
        for (int n = 0; n < q->N_taper; ++n) w[n] = 0.5f * (1.0f + std::cos(pi * (float)n / (float)q->N_taper));
    }
    const size_t nw = w.size(), nsf = q->nt_sf ? NumRec : 0, ncut = q->nt_cut ? NumRec : 0;
    void *va;
    if (int rc = get_buf(c, "cond_aux", (nw + nsf + ncut + 1) * 4, &va)) return rc;
    float *d_w = (float *)va;
    int *d_sf = (int *)va + nw, *d_cut = (int *)va + nw + nsf;
    if (nw) HIP_TRY(c, hipMemcpyAsync(d_w, w.data(), nw * 4, hipMemcpyHostToDevice, c->stream));
    if (nsf) HIP_TRY(c, hipMemcpyAsync(d_sf, q->nt_sf, nsf * 4, hipMemcpyHostToDevice, c->stream));
    if (ncut) HIP_TRY(c, hipMemcpyAsync(d_cut, q->nt_cut, ncut * 4, hipMemcpyHostToDevice, c->stream));
    hipExtLaunchKernelGGL(cond_resample_kernel, dim3((q->nt_taper + 255) / 256, NumRec), dim3(256), 0, c->stream,
                          timed ? c->cond_ev[2] : nullptr, timed ? c->cond_ev[3] : nullptr, 0, src, nt, q->dt_data,
                          resample ? 1 : 0, q->nt_prop2, q->dt_prop, q->nt_taper, mode, q->N_taper, d_w,
                          nsf ? d_sf : nullptr, ncut ? d_cut : nullptr, d_target);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (timed) {
        float ms;
        if (filt) {
            HIP_TRY(c, hipEventElapsedTime(&ms, c->cond_ev[0], c->cond_ev[1]));
            c->cond_ms[0] += ms;
            c->cond_n[0]++;
        }
        HIP_TRY(c, hipEventElapsedTime(&ms, c->cond_ev[2], c->cond_ev[3]));
        c->cond_ms[1] += ms;
        c->cond_n[1]++;
    }
    return 0;
}

extern "C" int psdm_rtm_condition_host(psdm_rtm_ctx *c, const psdm_rtm_cond *q, int NumRec, const float *field,
                                       float *target) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!q || !field || !target) return fail(c, PSDM_RTM_EINVAL, "condition: null argument");
    if (NumRec < 1 || q->nt_data < 1 || q->nt_taper < 1) return fail(c, PSDM_RTM_EINVAL, "condition: empty gather");
    if (int rc = set_device(c)) return rc;
    const size_t nin = (size_t)q->nt_data * NumRec, nout = (size_t)q->nt_taper * NumRec;
    void *vi, *vo;
    if (int rc = get_buf(c, "cond_in", nin * sizeof(float), &vi)) return rc;
    if (int rc = get_buf(c, "cond_out", nout * sizeof(float), &vo)) return rc;
    HIP_TRY(c, hipMemcpyAsync(vi, field, nin * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (int rc = psdm_rtm_condition(c, q, NumRec, (const float *)vi, (float *)vo)) return rc;
    HIP_TRY(c, hipMemcpyAsync(target, vo, nout * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int psdm_rtm_condition_timing(psdm_rtm_ctx *c, double *filter_ms, double *resample_ms, long long *n) {
    if (!c) return PSDM_RTM_EINVAL;
    if (filter_ms) *filter_ms = c->cond_n[0] ? c->cond_ms[0] / (double)c->cond_n[0] : 0.0;
    if (resample_ms) *resample_ms = c->cond_n[1] ? c->cond_ms[1] / (double)c->cond_n[1] : 0.0;
    if (n) *n = c->cond_n[1];
    return 0;
}
