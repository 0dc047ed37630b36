// Device helpers shared by the RTM step kernels (rtm_kernels.hip: one step per launch;
// rtm_tblock.hip: two steps per launch).
#pragma once
#include "rtm_core.hpp"

namespace psdm {

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
    return ((const float __attribute__((address_space(1))) *)a.c2)[clampi(iw, 0, a.nx - 1) * a.ldc + jw];
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


// Global-memory views of the step arguments: pointers read from a StepArgs in memory
// carry no address space of their own, and would otherwise be accessed with flat
// instructions (which wait on both the LDS and the vector-memory counters).
template <class T>
using gptr = T __attribute__((address_space(1))) *;
template <class T>
__device__ __forceinline__ gptr<T> gp(T *p) {
    return (gptr<T>)p;
}

// Wave-uniform read of memory that no launch in flight writes, through the scalar data
// cache (s_load): its latency and counter are separate from the vector loads.
template <class T>
__device__ __forceinline__ T sload(const T *p) {
    return *(const T __attribute__((address_space(4))) *)p;
}

// Native 4 x float / 2 x uint vectors (16- and 8-byte lane accesses).
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

// Lane L receives lane L-1's value (DPP wave_shr:1; lane 0 receives 0).
__device__ __forceinline__ float from_prev_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
// Lane L receives lane L+1's value (DPP wave_shl:1; lane 63 receives 0).
__device__ __forceinline__ float from_next_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// D2 of one row of the 6th-order stencil, in the reference's order (solver_space_derivs_clean.f90:1-69):
// ((((p+p)*c0 + (a1+b1)*c1) + (a2+b2)*c2) + (a3+b3)*c3)
__device__ __forceinline__ float d2(float p, float a1, float b1, float a2, float b2, float a3, float b3) {
    float d = (p + p) * kC2_0;
    d = d + (a1 + b1) * kC2_1;
    d = d + (a2 + b2) * kC2_2;
    d = d + (a3 + b3) * kC2_3;
    return d;
}

// The zero-sum form of the same stencil (opt-in, not the reference's arithmetic):
// sum over k of ((a_k - p) + (b_k - p)) * w_k, whose centre weight is exactly
// -2 (w_1 + w_2 + w_3), so a constant field has a zero derivative.  The FP32 weights of
// the reference form (p + p) * w_0 + ... sum to -5.9e-8 instead of 0: a spurious mass
// term whose phase error grows with the record (image rel-L2 against the FP64 oracle
// over 6000 steps of a 601 x 715 window: 1.1e-4 with d2, 2.1e-5 with d2z).
__device__ __forceinline__ float d2z(float p, float a1, float b1, float a2, float b2, float a3, float b3) {
    float d = ((a1 - p) + (b1 - p)) * kC2_1;
    d = d + ((a2 - p) + (b2 - p)) * kC2_2;
    d = d + ((a3 - p) + (b3 - p)) * kC2_3;
    return d;
}

// Lane address = wave-uniform base + this lane's 32-bit byte offset (global loads and
// stores then take the base in scalar registers and one offset register).
// (__builtin_amdgcn_readfirstlane is int -> int: the 64-bit base is read as two halves)
template <class T>
__device__ __forceinline__ gptr<T> lanep(const void *base, unsigned off) {
    const unsigned long long v = (unsigned long long)base;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return gp((T *)((const char *)(((unsigned long long)hi << 32) | lo) + off));
}
__device__ __forceinline__ v4f ldv(const float *base, unsigned off) { return *lanep<const v4f>(base, off); }

}  // namespace psdm
