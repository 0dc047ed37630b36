// The data formats on either side of the migration: SU trace files and the shot
// navigation, without Seismic Unix.
//
//   su_decode_kernel   SU traces (240-byte header + ns 4-byte samples each, big- or
//                      little-endian) -> FP32 gather (ntr, nt), trace-major: the first nt
//                      samples of each trace (get_shot_data get_mcs_data.f90:150-214,
//                      get_su_data auxiliar.f90:9-44), optionally shifted by the data delay
//                      (:184-204).  Pure byte movement, HBM-bound: the gather is one flat
//                      stream of output words, four per thread, lane-contiguous (coalesced).
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

constexpr int kSuHdr = 60;  // header words per trace (240 bytes, size_su_header)
constexpr int kSuBlock = 1024;  // output words per workgroup: 4 per thread, lane-contiguous
constexpr int kSuMaxNs = (2147483647 - 4 * kSuHdr) / 4;  // trace words and bytes stay int

__device__ __forceinline__ unsigned order(unsigned w, int big) { return big ? __builtin_bswap32(w) : w; }
__device__ __forceinline__ unsigned order16(unsigned h, int big) {  // a 16-bit field at the word's low address
    return big ? ((h & 0xffu) << 8 | (h >> 8 & 0xffu)) : (h & 0xffffu);
}

// Output word g of the (ntr, nt) gather, g = trace*nt + sample: the whole gather is one
// flat stream, so no workgroup idles on a trace's last partial chunk.  Each thread issues
// its four loads before any store; loads and stores are non-temporal (streamed once: 1-2 %
// faster than cached accesses on the 302 MB line).
__global__ __launch_bounds__(256) void su_decode_kernel(const unsigned *__restrict__ su, int ns, int nt, int big,
                                                        int shift, int total, FastDiv fd_nt,
                                                        float *__restrict__ out) {
    const int g0 = blockIdx.x * kSuBlock + threadIdx.x;
    unsigned v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int g = g0 + q * 256;
        const int j = fdiv(g, fd_nt), i = g - j * nt, k = i - shift;
        v[q] = (g < total && k >= 0 && k < nt) ? __builtin_nontemporal_load(su + (size_t)j * (kSuHdr + ns) + kSuHdr + k)
                                                : 0u;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int g = g0 + q * 256;
        if (g < total) __builtin_nontemporal_store(__uint_as_float(order(v[q], big)), out + g);
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

// Header word w (0..59) of trace j, as stored (SU segy.h byte offsets): tracl (bytes 0-3)
// and tracr (4-7) = j + 1, cdp (20-23), trid (28-29), ns (114-115), d1 (180-183), f1 = 0
// (184), d2 (188-191), f2 = 0 (192); every other byte 0.
__device__ unsigned header_word(int w, int j, int big, int trid, int nj, float di, float dj, double di_su) {
    switch (w) {
    case 0:
    case 1: return order((unsigned)(j + 1), big);
    case 5: {
        // suchw key1=cdp key2=tracl a=-di b=di: (a + b*tracl + c*key3)/d in double, c = 0,
        // d = 1, stored into the int field by truncation
        const double t = (double)(j + 1);
        return order((unsigned)(int)((-di_su + di_su * t + 0.0 * t) / 1.0), big);
    }
    case 7: return order16((unsigned)trid, big);
    case 28: return order16((unsigned)nj, big) << 16;
    case 45: return order(__float_as_uint(dj), big);
    case 47: return order(__float_as_uint(di), big);
    default: return 0u;
    }
}

// Output word g of the SU bytes, g = trace*(60 + nj) + word.
__global__ __launch_bounds__(256) void su_encode_kernel(const float *__restrict__ img, int nj, int big, int trid,
                                                        float di, float dj, double di_su, int total,
                                                        FastDiv fd_nw, unsigned *__restrict__ out) {
    const int g0 = blockIdx.x * kSuBlock + threadIdx.x, nw = kSuHdr + nj;
    unsigned v[4];
    int jw[4], ww[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int g = g0 + q * 256;
        jw[q] = fdiv(g, fd_nw);
        ww[q] = g - jw[q] * nw;
        v[q] = (g < total && ww[q] >= kSuHdr) ? __float_as_uint(img[(size_t)jw[q] * nj + ww[q] - kSuHdr]) : 0u;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int g = g0 + q * 256;
        if (g < total)
            out[g] = ww[q] >= kSuHdr ? order(v[q], big) : header_word(ww[q], jw[q], big, trid, nj, di, dj, di_su);
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
    if (ntr < 1 || ns < 1 || nt < 1 || nt > ns || ns > kSuMaxNs)
        return fail(c, PSDM_RTM_EINVAL, "su_decode: need ntr >= 1 and 1 <= nt <= ns < 2^29");
    if (nbytes < (long long)ntr * (kSuHdr + ns) * 4)
        return fail(c, PSDM_RTM_EINVAL, "su_decode: nbytes < ntr * (240 + 4*ns)");
    if ((long long)ntr * nt > 2147483647LL - 4 * kSuBlock)
        return fail(c, PSDM_RTM_EINVAL, "su_decode: ntr * nt must stay below 2^31 words");
    const int total = ntr * nt;
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (int q = 0; q < 2; ++q)
            if (!c->su_ev[q]) HIP_TRY(c, hipEventCreate(&c->su_ev[q]));
    hipExtLaunchKernelGGL(su_decode_kernel, dim3((total + kSuBlock - 1) / kSuBlock), dim3(256), 0, c->stream,
                          timed ? c->su_ev[0] : nullptr, timed ? c->su_ev[1] : nullptr, 0, (const unsigned *)d_su, ns,
                          nt, big_endian != 0, shift, total, make_fastdiv(nt), d_out);
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

extern "C" int psdm_rtm_su_read_host(psdm_rtm_ctx *c, const char *path, long long offset, int ntr, int ns, int nt,
                                     int big_endian, int shift, float *h_out) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!path || !h_out) return fail(c, PSDM_RTM_EINVAL, "su_read_host: null argument");
    if (ntr < 1 || ns < 1 || nt < 1 || nt > ns || ns > kSuMaxNs || offset < 0)
        return fail(c, PSDM_RTM_EINVAL, "su_read_host: need ntr >= 1, 1 <= nt <= ns < 2^29, offset >= 0");
    const size_t nbytes = (size_t)ntr * (kSuHdr + ns) * 4;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, PSDM_RTM_EINVAL, std::string("su_read_host: cannot open ") + path);
    if (int rc = set_device(c)) {
        std::fclose(f);
        return rc;
    }
    if (c->su_pinned_n < nbytes) {  // pinned staging, grown on demand, kept by the context
        if (c->su_pinned) (void)hipHostFree(c->su_pinned);
        c->su_pinned = nullptr;
        c->su_pinned_n = 0;
        if (hipHostMalloc(&c->su_pinned, nbytes, hipHostMallocDefault) != hipSuccess) {
            std::fclose(f);
            return psdm::host::hip_fail(c, hipErrorOutOfMemory, "hipHostMalloc(su staging)");
        }
        c->su_pinned_n = nbytes;
    }
    const bool ok = std::fseek(f, (long)offset, SEEK_SET) == 0 && std::fread(c->su_pinned, 1, nbytes, f) == nbytes;
    std::fclose(f);
    if (!ok) return fail(c, PSDM_RTM_EINVAL, std::string("su_read_host: ") + path + " is shorter than ntr traces");
    void *d_su, *d_out;
    if (int rc = psdm::host::get_buf(c, "su_bytes", nbytes, &d_su)) return rc;
    if (int rc = psdm::host::get_buf(c, "su_gather", (size_t)ntr * nt * sizeof(float), &d_out)) return rc;
    HIP_TRY(c, hipMemcpyAsync(d_su, c->su_pinned, nbytes, hipMemcpyHostToDevice, c->stream));
    if (int rc = psdm_rtm_su_decode(c, d_su, (long long)nbytes, ntr, ns, nt, big_endian, shift, (float *)d_out))
        return rc;
    HIP_TRY(c, hipMemcpy(h_out, d_out, (size_t)ntr * nt * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int psdm_rtm_su_header_i32(psdm_rtm_ctx *c, const void *d_su, long long nbytes, int ntr, int ns,
                                      int big_endian, int pos, int *d_out) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_su || !d_out) return fail(c, PSDM_RTM_EINVAL, "su_header_i32: null argument");
    if (ntr < 1 || ns < 1 || ns > kSuMaxNs)
        return fail(c, PSDM_RTM_EINVAL, "su_header_i32: need ntr >= 1, 1 <= ns < 2^29");
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
    if ((long long)ni * (kSuHdr + nj) > 2147483647LL - 4 * kSuBlock)
        return fail(c, PSDM_RTM_EINVAL, "su_encode: ni * (60 + nj) must stay below 2^31 words");
    const int total = ni * (kSuHdr + nj);
    if (int rc = set_device(c)) return rc;
    const bool timed = c->timing != 0;
    if (timed)
        for (int q = 2; q < 4; ++q)
            if (!c->su_ev[q]) HIP_TRY(c, hipEventCreate(&c->su_ev[q]));
    hipExtLaunchKernelGGL(su_encode_kernel, dim3((total + kSuBlock - 1) / kSuBlock), dim3(256), 0, c->stream,
                          timed ? c->su_ev[2] : nullptr, timed ? c->su_ev[3] : nullptr, 0, d_img, nj, big_endian != 0,
                          trid, di, dj, su_parsed(di), total, make_fastdiv(kSuHdr + nj), (unsigned *)d_su);
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
