"""The multiply-shift division the kernels use for tile / slot / snapshot indexing
(psdm_amd/csrc/fastdiv.hpp) equals integer division: brute force over divisors
1..3000 and dividends 0..2^20 plus edge values up to 2^31 - 1, compiled with g++."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include <initializer_list>
#include "fastdiv.hpp"
int main() {
    using namespace psdm;
    long long bad = 0, n = 0;
    const int edges[] = {0, 1, 2, 3, 63, 64, 65, 1023, 1024, 4095, 65535, 65536, 1 << 24, (1 << 24) + 1,
                         1 << 30, (1 << 30) + 7, 2147483646, 2147483647};
    for (int d = 1; d <= 3000; ++d) {
        const FastDiv f = make_fastdiv(d);
        for (int x = 0; x < (1 << 20); x += (d < 64 ? 1 : 7)) { n++; if (fdiv(x, f) != x / d) bad++; }
        for (int e : edges) for (int o = -2; o <= 2; ++o) {
            long long x = (long long)e + o;
            if (x < 0 || x > 2147483647LL) continue;
            n++; if (fdiv((int)x, f) != (int)(x / d)) bad++;
        }
    }
    for (int d : {65536, 99991, 1 << 20, 123456789, 2147483647}) {
        const FastDiv f = make_fastdiv(d);
        for (long long x = 0; x <= 2147483647LL; x += 9973) { n++; if (fdiv((int)x, f) != (int)(x / d)) bad++; }
    }
    std::printf("%lld %lld\n", n, bad);
    return bad != 0;
}
"""


def test_fastdiv_matches_integer_division():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "t.cpp"), os.path.join(d, "t")
        open(src, "w").write(PROG)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "psdm_amd", "csrc"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        n, bad = map(int, out.stdout.split())
        assert out.returncode == 0 and bad == 0 and n > 10_000_000
