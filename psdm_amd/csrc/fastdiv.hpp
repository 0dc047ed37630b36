// Division by a divisor fixed for many launches, as a multiply-high and two shifts
// (Granlund & Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994, fig. 4.1): for 1 <= d < 2^31 and 0 <= x < 2^31,
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   t = (m * x) >> 32,  x / d = (t + ((x - t) >> s1)) >> s2  with s1 = min(l, 1), s2 = max(l - 1, 0).
// The multiplier is built on the host; on the device the quotient costs a handful of
// scalar instructions instead of an integer-division sequence.
#pragma once

#if defined(__HIPCC__)
#define PSDM_HD __host__ __device__
#else
#define PSDM_HD
#endif

namespace psdm {

struct FastDiv {
    unsigned m;
    int s1, s2, d;
};

inline FastDiv make_fastdiv(int d) {
    int l = 0;
    while ((1ull << l) < (unsigned long long)d) ++l;
    FastDiv f;
    f.m = (unsigned)((((1ull << 32) * ((1ull << l) - (unsigned long long)d)) / (unsigned long long)d) + 1ull);
    f.s1 = l > 0 ? 1 : 0;
    f.s2 = l > 0 ? l - 1 : 0;
    f.d = d;
    return f;
}

PSDM_HD inline int fdiv(int x, const FastDiv &f) {
    const unsigned t = (unsigned)(((unsigned long long)f.m * (unsigned)x) >> 32);
    return (int)((t + (((unsigned)x - t) >> f.s1)) >> f.s2);
}

}  // namespace psdm
