// The data formats on either side of the migration: SU trace files and the shot
// navigation, without Seismic Unix.
//
//   su_decode_kernel   SU traces (240-byte header + ns 4-byte samples each, big- or
//                      little-endian) -> FP32 gather (ntr, nt), trace-major: the first nt
//                      samples of each trace (get_shot_data get_mcs_data.f90:150-214,
//                      get_su_data auxiliar.f90:9-44), optionally shifted by the data delay
//                      (:184-204).  Pure byte movement: HBM-bound, one 4-byte word per thread
//                      per step, 1024 samples of one trace per workgroup, coalesced.
//   su_header_kernel   one INTEGER*4 header word per trace at a 1-based byte position, as
//                      get_pos_channels reads the shot number and offsets (:449-490).
//   su_encode_kernel   FP32 image (ni traces of nj samples) -> SU bytes with the header
//                      the reference's Seismic Unix pipeline writes (write_psdm
//                      auxiliar.f90:275-299: suaddhead ns=nj, sushw d1/f1/d2/f2/trid,
//                      suchw cdp).
//
// Host functions (no device work): the receiver and source grid indices of a shot
// (get_pos_channels :418-510, get_geometry_mcs :128-148) and the navigation file reader
// (read_nav_shot_mcs :30-126).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fastdiv.hpp"
#include "rtm_ctx.hpp"

using psdm::FastDiv;
using psdm::fdiv;
using psdm::make_fastdiv;
using psdm::host::fail;
using psdm::host::set_device;

namespace psdm {
namespace {

constexpr int kSuHdr = 60;      // header words per trace (240 bytes, size_su_header)
constexpr int kSuChunk = 1024;  // samples (words) of one trace per workgroup

__device__ __forceinline__ unsigned order(unsigned w, int big) { return big ? __builtin_bswap32(w) : w; }

__global__ __launch_bounds__(256) void su_decode_kernel(const unsigned *__restrict__ su, int ns, int nt, int big,
                                                        int shift, int nchunk, FastDiv fd_chunk,
                                                        float *__restrict__ out) {
    const int b = blockIdx.x, j = fdiv(b, fd_chunk), ch = b - j * nchunk;
    const unsigned *tr = su + (size_t)j * (kSuHdr + ns) + kSuHdr;
    float *o = out + (size_t)j * nt;
    const int i0 = ch * kSuChunk + threadIdx.x;
    unsigned v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // all loads first: four independent words in flight per thread
        const int k = i0 + q * 256 - shift;
        v[q] = (k >= 0 && k < nt && i0 + q * 256 < nt) ? tr[k] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + q * 256;
        if (i < nt) o[i] = __uint_as_float(order(v[q], big));
    }
}

__global__ void su_header_kernel(const unsigned char *__restrict__ su, int ns, int ntr, int pos0, int big,
                                 int *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntr) return;
    const unsigned char *p = su + (size_t)j * (kSuHdr + ns) * 4 + pos0;
    const unsigned v = big ? ((unsigned)p[0] << 24 | (unsigned)p[1] << 16 | (unsigned)p[2] << 8 | p[3])
                           : ((unsigned)p[3] << 24 | (unsigned)p[2] << 16 | (unsigned)p[1] << 8 | p[0]);
    out[j] = (int)v;
}

// Byte `off` (0..239) of the header of trace j.  Fields (SU segy.h byte offsets):
// tracl 0 and tracr 4 (int32, = j + 1), cdp 20 (int32), trid 28 (int16), ns 114 (uint16),
// d1 180, f1 184, d2 188, f2 192 (float); every other byte 0.
struct HdrVals {
    unsigned tracl, cdp, trid, ns, d1, f1, d2, f2;
};
__device__ __forceinline__ unsigned field_byte(int off, int start, int size, unsigned v, int big) {
    const int k = off - start;
    return (unsigned)(v >> (big ? 8 * (size - 1 - k) : 8 * k)) & 0xffu;
}
__device__ unsigned hdr_byte(int off, const HdrVals &h, int big) {
    if (off < 4) return field_byte(off, 0, 4, h.tracl, big);
    if (off < 8) return field_byte(off, 4, 4, h.tracl, big);
    if (off >= 20 && off < 24) return field_byte(off, 20, 4, h.cdp, big);
    if (off >= 28 && off < 30) return field_byte(off, 28, 2, h.trid, big);
    if (off >= 114 && off < 116) return field_byte(off, 114, 2, h.ns, big);
    if (off >= 180 && off < 196) {
        const int f = (off - 180) >> 2;
        const unsigned v = f == 0 ? h.d1 : f == 1 ? h.f1 : f == 2 ? h.d2 : h.f2;
        return field_byte(off, 180 + 4 * f, 4, v, big);
    }
    return 0u;
}

__global__ __launch_bounds__(256) void su_encode_kernel(const float *__restrict__ img, int nj, int big, int trid,
                                                        float di, float dj, double di_su, int nchunk,
                                                        FastDiv fd_chunk, unsigned *__restrict__ out) {
    const int b = blockIdx.x, j = fdiv(b, fd_chunk), ch = b - j * nchunk;
    const int nw = kSuHdr + nj;
    unsigned *o = out + (size_t)j * nw;
    const float *src = img + (size_t)j * nj;
    const int w0 = ch * kSuChunk + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int w = w0 + q * 256;
        if (w >= nw) break;
        if (w >= kSuHdr) {
            o[w] = order(__float_as_uint(src[w - kSuHdr]), big);
        } else {
            // suchw key1=cdp key2=tracl a=-di b=di: (a + b*tracl + c*key3)/d in double,
            // c = 0, d = 1, stored into the int field by truncation
            const double cd = (-di_su + di_su * (double)(j + 1) + 0.0 * (double)(j + 1)) / 1.0;
            const HdrVals h{(unsigned)(j + 1), (unsigned)(int)cd, (unsigned)trid & 0xffffu, (unsigned)nj & 0xffffu,
                            __float_as_uint(dj), 0u, __float_as_uint(di), 0u};
            unsigned v = 0;
            for (int k = 0; k < 4; ++k) v |= hdr_byte(4 * w + k, h, big) << (8 * k);
            o[w] = v;
        }
    }
}

// The value Seismic Unix parses for a real written by the reference's list-directed
// `write(Str,*) x` (9 significant digits for a REAL(4)).
double su_parsed(float x) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.8E", (double)x);
    return std::strtod(buf, nullptr);
}

bool aligned4(const void *p) { return ((uintptr_t)p & 3u) == 0; }

}  // namespace
}  // namespace psdm

using namespace psdm;

extern "C" int psdm_rtm_su_decode(psdm_rtm_ctx *c, const void *d_su, long long nbytes, int ntr, int ns, int nt,
                                  int big_endian, int shift, float *d_out) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_su || !d_out) return fail(c, PSDM_RTM_EINVAL, "su_decode: null argument");
    if (!aligned4(d_su)) return fail(c, PSDM_RTM_EINVAL, "su_decode: the SU bytes must be 4-byte aligned");
    if (ntr < 1 || ns < 1 || nt < 1 || nt > ns)
        return fail(c, PSDM_RTM_EINVAL, "su_decode: need ntr >= 1 and 1 <= nt <= ns");
    if (nbytes < (long long)ntr * (kSuHdr + ns) * 4)
        return fail(c, PSDM_RTM_EINVAL, "su_decode: nbytes < ntr * (240 + 4*ns)");
    const int nchunk = (nt + kSuChunk - 1) / kSuChunk;
    if ((long long)ntr * nchunk > 2147483647LL) return fail(c, PSDM_RTM_EINVAL, "su_decode: too many traces");
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (int q = 0; q < 2; ++q)
            if (!c->su_ev[q]) HIP_TRY(c, hipEventCreate(&c->su_ev[q]));
    hipExtLaunchKernelGGL(su_decode_kernel, dim3(ntr * nchunk), dim3(256), 0, c->stream,
                          timed ? c->su_ev[0] : nullptr, timed ? c->su_ev[1] : nullptr, 0, (const unsigned *)d_su, ns,
                          nt, big_endian != 0, shift, nchunk, make_fastdiv(nchunk), d_out);
    HIP_TRY(c, hipGetLastError());
    if (timed) {
        HIP_TRY(c, hipEventSynchronize(c->su_ev[1]));
        float ms;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->su_ev[0], c->su_ev[1]));
        c->su_ms[0] += ms;
        c->su_n[0]++;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int psdm_rtm_su_header_i32(psdm_rtm_ctx *c, const void *d_su, long long nbytes, int ntr, int ns,
                                      int big_endian, int pos, int *d_out) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_su || !d_out) return fail(c, PSDM_RTM_EINVAL, "su_header_i32: null argument");
    if (ntr < 1 || ns < 1) return fail(c, PSDM_RTM_EINVAL, "su_header_i32: need ntr >= 1, ns >= 1");
    if (pos < 1 || pos > (kSuHdr + ns) * 4 - 3)
        return fail(c, PSDM_RTM_EINVAL, "su_header_i32: byte position outside the trace");
    if (nbytes < (long long)ntr * (kSuHdr + ns) * 4)
        return fail(c, PSDM_RTM_EINVAL, "su_header_i32: nbytes < ntr * (240 + 4*ns)");
    if (int rc = set_device(c)) return rc;
    hipLaunchKernelGGL(su_header_kernel, dim3((ntr + 255) / 256), dim3(256), 0, c->stream,
                       (const unsigned char *)d_su, ns, ntr, pos - 1, big_endian != 0, d_out);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int psdm_rtm_su_encode(psdm_rtm_ctx *c, const float *d_img, int ni, int nj, float di, float dj, int trid,
                                  int big_endian, void *d_su) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_img || !d_su) return fail(c, PSDM_RTM_EINVAL, "su_encode: null argument");
    if (!aligned4(d_su)) return fail(c, PSDM_RTM_EINVAL, "su_encode: the output must be 4-byte aligned");
    if (ni < 1 || nj < 1 || nj > 65535)
        return fail(c, PSDM_RTM_EINVAL, "su_encode: need ni >= 1 and 1 <= nj <= 65535 (ns is 16 bits)");
    if (trid < -32768 || trid > 32767) return fail(c, PSDM_RTM_EINVAL, "su_encode: trid is a 16-bit field");
    const int nchunk = (kSuHdr + nj + kSuChunk - 1) / kSuChunk;
    if ((long long)ni * nchunk > 2147483647LL) return fail(c, PSDM_RTM_EINVAL, "su_encode: too many traces");
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (int q = 2; q < 4; ++q)
            if (!c->su_ev[q]) HIP_TRY(c, hipEventCreate(&c->su_ev[q]));
    hipExtLaunchKernelGGL(su_encode_kernel, dim3(ni * nchunk), dim3(256), 0, c->stream,
                          timed ? c->su_ev[2] : nullptr, timed ? c->su_ev[3] : nullptr, 0, d_img, nj, big_endian != 0,
                          trid, di, dj, su_parsed(di), nchunk, make_fastdiv(nchunk), (unsigned *)d_su);
    HIP_TRY(c, hipGetLastError());
    if (timed) {
        HIP_TRY(c, hipEventSynchronize(c->su_ev[3]));
        float ms;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->su_ev[2], c->su_ev[3]));
        c->su_ms[1] += ms;
        c->su_n[1]++;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int psdm_rtm_su_timing(psdm_rtm_ctx *c, double *decode_ms, double *encode_ms, long long *n_decode,
                                  long long *n_encode) {
    if (!c) return PSDM_RTM_EINVAL;
    if (decode_ms) *decode_ms = c->su_n[0] ? c->su_ms[0] / (double)c->su_n[0] : 0.0;
    if (encode_ms) *encode_ms = c->su_n[1] ? c->su_ms[1] / (double)c->su_n[1] : 0.0;
    if (n_decode) *n_decode = c->su_n[0];
    if (n_encode) *n_encode = c->su_n[1];
    return 0;
}

extern "C" int psdm_rtm_mcs_geometry(const int *offsets, int nrec, float pos_shot, int offset_header,
                                     float offset_unit, int dir_boat, float near_offset, float drec, float dd,
                                     float shot_depth, float streamer_depth, float *pos_rec, int *nxRec, int *nyRec,
                                     int *nxSou, int *nySou) {
    if (nrec < 0 || !(dd > 0.f) || !nxSou || !nySou || (nrec > 0 && (!pos_rec || !nxRec || !nyRec)))
        return PSDM_RTM_EINVAL;
    if (offset_header == 1 && (!offsets || offset_unit == 0.f)) return PSDM_RTM_EINVAL;
    if (offset_header != 0 && offset_header != 1) return PSDM_RTM_EINVAL;
    const float dirf = (float)dir_boat;
    for (int r = 0; r < nrec; ++r) {
        float off;
        if (offset_header == 1)  // offset_su = |offset| * unit, or |offset| / |unit| for unit < 0
            off = offset_unit > 0.f ? (float)std::abs(offsets[r]) * offset_unit
                                    : (float)std::abs(offsets[r]) / std::fabs(offset_unit);
        else
            off = near_offset + (float)r * drec;
        const float prod = dirf * off;
        pos_rec[r] = pos_shot - prod;
        const float q = pos_rec[r] / dd;
        nxRec[r] = 1 + (int)std::floor(q);
        const float qy = streamer_depth / dd;
        nyRec[r] = 1 + (int)std::floor(qy);
    }
    const float qs = pos_shot / dd, qd = shot_depth / dd;
    *nxSou = 1 + (int)std::floor(qs);
    *nySou = 1 + (int)std::floor(qd);
    return 0;
}

extern "C" int psdm_rtm_nav_read(const char *path, int cols, int dir_boat, float d_shot_model, int *shot_i,
                                 int *shot_f, int cap, int *ids, float *pos, float *bat, int *n, int *line_i,
                                 int *line_f) {
    if (!path || !shot_i || !shot_f || !n || !line_i || !line_f) return PSDM_RTM_EINVAL;
    if (cols != 11 && cols != 111 && cols != 1111) return PSDM_RTM_EINVAL;
    if (dir_boat != 1 && dir_boat != -1) return PSDM_RTM_EINVAL;
    FILE *f = std::fopen(path, "r");
    if (!f) return PSDM_RTM_EINVAL;
    std::vector<int> id;
    std::vector<float> xs, ys, bt;
    std::string line;
    int ch, bad = 0;
    do {  // one record per line; a last line without a newline counts
        line.clear();
        while ((ch = std::fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
        if (ch == EOF && line.empty()) break;
        for (char &q : line)
            if (q == ',' || q == '\t' || q == '\r') q = ' ';
        const char *p = line.c_str();
        char *e;
        errno = 0;
        const long v = std::strtol(p, &e, 10);
        if (e == p || errno || (*e != ' ' && *e != '\0')) bad = 1;
        p = e;
        float val[3] = {0.f, 0.f, 0.f};
        const int nv = cols == 11 ? 1 : cols == 111 ? 2 : 3;
        for (int k = 0; k < nv; ++k) {
            val[k] = std::strtof(p, &e);
            if (e == p) bad = 1;
            p = e;
        }
        id.push_back((int)v);
        xs.push_back(val[0]);
        ys.push_back(val[1]);
        bt.push_back(val[2]);
    } while (ch != EOF);
    std::fclose(f);
    if (bad) return PSDM_RTM_EINVAL;
    const int nl = (int)id.size();
    *n = nl;
    if (nl == 0) return PSDM_RTM_EINVAL;
    if (nl > cap || !ids || !pos || !bat) return cap == 0 ? 0 : PSDM_RTM_EINVAL;
    const int r0 = dir_boat == 1 ? 0 : nl - 1;  // the line's origin: first shot sailing +x, last sailing -x
    const float x1 = xs[r0], y1 = ys[r0];
    *line_i = 0;
    *line_f = 0;
    for (int r = 0; r < nl; ++r) {
        if (r == 0 && *shot_i == 0) *shot_i = id[r];
        if (r == nl - 1 && *shot_f == 0) *shot_f = id[r];
        if (*shot_i > 0 && *shot_i == id[r]) *line_i = r + 1;
        if (*shot_f > 0 && *shot_f == id[r]) *line_f = r + 1;
    }
    for (int r = 0; r < nl; ++r) {
        const float dx = xs[r] - x1, dy = ys[r] - y1;
        const float sx = dx * dx, sy = dy * dy;
        const float s = sx + sy;
        ids[r] = id[r];
        pos[r] = d_shot_model + std::sqrt(s);
        bat[r] = std::fabs(bt[r]);
    }
    return 0;
}
