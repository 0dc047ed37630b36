"""The SU / navigation oracle (oracle/su_codec.py) pinned against the reference's own
reader, geometry and navigation routines compiled into oracle/_ref (open_shot_mcs +
get_shot_data, get_pos_channels + get_geometry_mcs, read_nav_shot_mcs, get_su_data), and
the SU writer against the trace-header table of Seismic Unix's segy.h."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from oracle import pyoracle
from oracle import su_codec as S

need_ref = pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built")


def su_file(tmp_path, rng, ntr, ns, big, name="shot.su", special=False):
    """An SU file of random headers and samples; returns (path, bytes, samples as stored)."""
    e = ">" if big else "<"
    hdr = rng.integers(-2**31, 2**31 - 1, (ntr, 60)).astype(e + "i4")
    smp = rng.standard_normal((ntr, ns)).astype(np.float32)
    if special and ns >= 8:  # -0, infinities, a denormal, NaNs with payloads
        bits = np.array([0x80000000, 0x7F800000, 0xFF800000, 0x00000001, 0x7FC00001, 0xFFA5A5A5, 0x7F7FFFFF,
                         0x00800000], np.uint32)
        smp[0, :8] = bits.view(np.float32)
    buf = b"".join(hdr[j].tobytes() + smp[j].astype(e + "f4").tobytes() for j in range(ntr))
    p = os.path.join(str(tmp_path), name)
    with open(p, "wb") as f:
        f.write(buf)
    return p, buf, smp


@need_ref
@pytest.mark.parametrize("ntr,ns,nt,big", [(7, 50, 43, True), (7, 50, 50, False), (1, 1, 1, True),
                                           (480, 2401, 2401, True), (33, 1000, 601, False)])
def test_decode_bit_exact_vs_reference_reader(tmp_path, ntr, ns, nt, big):
    rng = np.random.default_rng(ntr * 7 + ns)
    p, buf, smp = su_file(tmp_path, rng, ntr, ns, big, special=True)
    a = S.decode(buf, ntr, ns, nt, big)
    b = S.ref_read_shot(p, big, ntr, ns, nt)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(a.view(np.uint32), smp[:, :nt].view(np.uint32))


def test_decode_shift_invariants():
    """The delayed gather (get_mcs_data.f90:184-204): a shift by s then by -s restores the
    samples that stayed inside; |s| >= nt leaves zeros; s = 0 is the raw gather."""
    rng = np.random.default_rng(3)
    ntr, ns, nt = 5, 40, 30
    buf = b"".join(rng.integers(0, 255, 240, dtype=np.uint8).tobytes()
                   + rng.standard_normal(ns).astype(">f4").tobytes() for _ in range(ntr))
    raw = S.decode(buf, ntr, ns, nt)
    for s in (1, 7, 29, -1, -12, -29):
        d = S.decode(buf, ntr, ns, nt, shift=s)
        if s > 0:
            np.testing.assert_array_equal(d[:, s:], raw[:, :nt - s])
            assert not d[:, :s].any()
        else:
            np.testing.assert_array_equal(d[:, :nt + s], raw[:, -s:])
            assert not d[:, nt + s:].any()
    for s in (30, 31, -30, -100):
        assert not S.decode(buf, ntr, ns, nt, shift=s).any()


@need_ref
@pytest.mark.parametrize("case", range(8))
def test_header_words_and_geometry_vs_reference(tmp_path, case):
    """Shot number (byte 9) and offsets (byte 37) from the headers, and the receiver and
    source grid indices, for both offset units, both sailing directions and the
    offset-free geometry (near_offset + r*drec)."""
    rng = np.random.default_rng(100 + case)
    big = case % 2 == 0
    ntr, ns = 48, 64
    e = ">" if big else "<"
    hdr = np.zeros((ntr, 60), e + "i4")
    hdr[:, 2] = 1234 + case  # fldr (byte 9)
    offs = np.round(150 + 12.5 * np.arange(ntr) * (1 if case % 3 else -1)).astype(np.int64)
    if case >= 4:
        offs = offs * 10  # decimetres
    hdr[:, 9] = offs  # offset (byte 37)
    buf = b"".join(hdr[j].tobytes() + np.zeros(ns, e + "f4").tobytes() for j in range(ntr))
    p = os.path.join(str(tmp_path), "g.su")
    with open(p, "wb") as f:
        f.write(buf)
    oh = 0 if case == 3 else 1
    unit = (1.0, 0.5, 2.0, 1.0, -10.0, -4.0, 0.1, -10.0)[case]
    dirb = 1 if case % 4 < 2 else -1
    dd = (10.0, 12.5, 6.25, 10.0, 10.0, 7.5, 10.0, 2.5)[case]
    pos_shot = float(np.float32(rng.uniform(6000, 9000)))
    if case == 0:
        pos_shot = 8000.0  # receivers on grid lines: floor at exact multiples
    kw = dict(offset_header=oh, offset_unit=unit, dir_boat=dirb, near_offset=150.0, drec=12.5, dd=dd,
              shot_depth=5.0 + case, streamer_depth=7.5 + case)
    sid, offs_su, posr, nxr, nyr, nxs, nys = S.ref_geometry(p, big, ntr, ns, 9, oh, unit, dirb, 150.0, 12.5,
                                                           pos_shot, dd, 5.0 + case, 7.5 + case)
    assert sid == S.header_i32(buf, 1, ns, 9, big)[0] == 1234 + case
    h = S.header_i32(buf, ntr, ns, 37, big)
    np.testing.assert_array_equal(h, offs)
    g = S.geometry(h, ntr, pos_shot, **kw)
    np.testing.assert_array_equal(g[0], posr)
    np.testing.assert_array_equal(g[1], nxr)
    np.testing.assert_array_equal(g[2], nyr)
    assert (g[3], g[4]) == (nxs, nys)


@need_ref
@pytest.mark.parametrize("cols,dirb", [(11, 1), (111, 1), (1111, -1), (1111, 1), (111, -1)])
def test_navigation_vs_reference(tmp_path, cols, dirb):
    rng = np.random.default_rng(cols + dirb)
    n = 57
    ids = 1001 + np.arange(n)
    x = 500000.0 + np.cumsum(rng.uniform(24, 26, n))
    y = 4200000.5 + np.cumsum(rng.uniform(-3, 3, n))
    z = -rng.uniform(100, 3000, n)
    p = os.path.join(str(tmp_path), "nav.txt")
    with open(p, "w") as f:
        for r in range(n):
            vals = [f"{x[r]:.2f}", f"{y[r]:.3f}", f"{z[r]:.1f}"][: {11: 1, 111: 2, 1111: 3}[cols]]
            f.write(f"{ids[r]} " + " ".join(vals) + "\n")
    a = S.read_nav(p, cols, dirb, 1500.0)
    b = S.ref_read_nav(p, cols, dirb, 1500.0, n)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    # pos = add1 + sqrt((xs-x1)**2. + (ys-y1)**2.) (get_mcs_data.f90:110).  The reference's
    # own build (gfortran -O3, src/Makefile:3) folds x**2.0 into the correctly rounded x*x;
    # oracle/_ref compiles this file at -O0, where amdflang calls powf, and glibc's powf
    # rounds an exact halfway square the other way.  Both agree wherever no square is a tie.
    libm = C.CDLL("libm.so.6")
    libm.powf.restype, libm.powf.argtypes = C.c_float, [C.c_float, C.c_float]
    xs = np.array([np.float32(float(v)) for v in np.loadtxt(p, usecols=1, ndmin=1)], np.float32)
    ys = (np.array([np.float32(v) for v in np.loadtxt(p, usecols=2, ndmin=1)], np.float32) if cols > 11
          else np.zeros(n, np.float32))
    r0 = 0 if dirb == 1 else n - 1
    tie = np.zeros(n, bool)
    for r in range(n):
        for d in (np.float32(xs[r] - xs[r0]), np.float32(ys[r] - ys[r0])):
            tie[r] |= np.float32(d * d) != np.float32(libm.powf(float(d), 2.0))
    np.testing.assert_array_equal(a[1][~tie].view(np.uint32), b[1][~tie].view(np.uint32))
    if cols == 1111:
        assert tie.any()  # the comparison above covers the other cases; make sure ties occur
    assert (a[5], a[6]) == (b[3], b[4]) == (1, n)
    assert (a[3], a[4]) == (ids[0], ids[-1])


# Seismic Unix segy.h trace header: (name, byte offset, struct code)
SEGY = [("tracl", 0, "i"), ("tracr", 4, "i"), ("fldr", 8, "i"), ("tracf", 12, "i"), ("ep", 16, "i"),
        ("cdp", 20, "i"), ("cdpt", 24, "i"), ("trid", 28, "h"), ("nvs", 30, "h"), ("nhs", 32, "h"),
        ("duse", 34, "h"), ("offset", 36, "i"), ("ns", 114, "H"), ("dt", 116, "H"), ("d1", 180, "f"),
        ("f1", 184, "f"), ("d2", 188, "f"), ("f2", 192, "f"), ("ungpow", 196, "f"), ("unscale", 200, "f"),
        ("ntr", 204, "i"), ("mark", 208, "h")]


@pytest.mark.parametrize("big", [True, False])
def test_encode_header_table(big):
    rng = np.random.default_rng(5)
    ni, nj = 6, 11
    img = rng.standard_normal((ni, nj)).astype(np.float32)
    di, dj = 12.5, 10.0
    b = S.encode(img, di, dj, trid=130, big=big)
    assert len(b) == ni * (240 + 4 * nj)
    e = ">" if big else "<"
    for j in range(ni):
        h = b[j * (240 + 4 * nj): j * (240 + 4 * nj) + 240]
        v = {nm: struct.unpack_from(e + code, h, off)[0] for nm, off, code in SEGY}
        assert v["tracl"] == v["tracr"] == j + 1
        assert v["cdp"] == int(12.5 * j)  # suchw key1=cdp key2=tracl a=-di b=di
        assert v["trid"] == 130 and v["ns"] == nj
        assert (v["d1"], v["f1"], v["d2"], v["f2"]) == (10.0, 0.0, 12.5, 0.0)
        named = {"tracl", "tracr", "cdp", "trid", "ns", "d1", "f1", "d2", "f2"}
        assert all(v[k] == 0 for k in v if k not in named)
        mask = np.ones(240, bool)
        for nm, off, code in SEGY:
            if nm in named:
                mask[off:off + struct.calcsize(code)] = False
        assert not np.frombuffer(h, np.uint8)[mask].any()


def test_encode_cdp_follows_the_parsed_spacing():
    """suchw computes cdp in double from the 9-digit text of di: 0.1 becomes 0.100000001."""
    assert S.su_parsed(0.1) == 0.100000001
    b = S.encode(np.zeros((12, 3), np.float32), 0.1, 1.0, trid=1)
    cdp = [struct.unpack_from(">i", b, j * 252 + 20)[0] for j in range(12)]
    assert cdp == [int(-0.100000001 + 0.100000001 * (j + 1)) for j in range(12)]


@need_ref
@pytest.mark.parametrize("big", [True, False])
def test_encode_read_back_by_reference_reader(tmp_path, big):
    """Image -> SU file -> the reference's get_su_data (auxiliar.f90:9-44) and its MCS
    reader: the samples come back bit for bit."""
    rng = np.random.default_rng(9)
    ni, nj = 25, 301
    img = rng.standard_normal((ni, nj)).astype(np.float32)
    b = S.encode(img, 10.0, 10.0, trid=1, big=big)
    p = os.path.join(str(tmp_path), "img.su")
    with open(p, "wb") as f:
        f.write(b)
    np.testing.assert_array_equal(S.ref_get_su_data(p, big, ni, nj, nj).view(np.uint32), img.view(np.uint32))
    np.testing.assert_array_equal(S.ref_get_su_data(p, big, ni, nj, 200), img[:, :200])
    np.testing.assert_array_equal(S.ref_read_shot(p, big, ni, nj, nj), img)
    np.testing.assert_array_equal(S.decode(b, ni, nj, nj, big), img)
