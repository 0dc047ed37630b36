"""The library's host-side SU geometry and navigation reader (psdm_rtm_mcs_geometry,
psdm_rtm_nav_read: no device work) against the oracle (oracle/su_codec.py, itself pinned
against the reference's routines in tests/test_su_oracle.py)."""
import os

import numpy as np
import pytest

from oracle import su_codec as S


@pytest.fixture(scope="module")
def su():
    from psdm_amd import _native, su

    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libpsdm_rtm.so not built")
    return su


@pytest.mark.parametrize("case", range(10))
def test_geometry_matches_oracle(su, case):
    rng = np.random.default_rng(case)
    nrec = int(rng.integers(1, 600))
    offs = rng.integers(-8000, 8000, nrec).astype(np.int32)
    kw = dict(offset_header=int(case % 3 != 2), offset_unit=float((1.0, -10.0, 0.25, 3.0, -0.5)[case % 5]),
              dir_boat=(1, -1)[case % 2], near_offset=float(np.float32(rng.uniform(0, 300))),
              drec=float(np.float32(rng.choice([6.25, 12.5, 25.0]))), dd=float((10.0, 12.5, 2.5, 6.25, 7.0)[case % 5]),
              shot_depth=float(np.float32(rng.uniform(0, 20))), streamer_depth=float(np.float32(rng.uniform(0, 30))))
    pos_shot = float(np.float32(rng.uniform(8000, 40000)))
    a = su.mcs_geometry(offs, nrec, pos_shot, **kw)
    b = S.geometry(offs, nrec, pos_shot, **kw)
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    assert (a[3], a[4]) == (b[3], b[4])


def test_geometry_rejects_bad_arguments(su):
    with pytest.raises(ValueError):
        su.mcs_geometry(np.zeros(3, np.int32), 3, 100.0, offset_header=1, offset_unit=0.0)
    with pytest.raises(ValueError):
        su.mcs_geometry(None, 3, 100.0, offset_header=1)
    with pytest.raises(ValueError):
        su.mcs_geometry(None, 3, 100.0, dd=0.0)


@pytest.mark.parametrize("cols,dirb,sep", [(11, 1, " "), (111, -1, ","), (1111, 1, "\t"), (1111, -1, "  ")])
def test_navigation_matches_oracle(su, tmp_path, cols, dirb, sep):
    rng = np.random.default_rng(cols)
    n = 300
    x = 431000.0 + np.cumsum(rng.uniform(20, 30, n))
    y = 4750000.0 + np.cumsum(rng.uniform(-5, 5, n))
    z = -rng.uniform(50, 4000, n)
    p = os.path.join(str(tmp_path), "nav.txt")
    with open(p, "w") as f:
        for r in range(n):
            vals = [f"{x[r]:.2f}", f"{y[r]:.2f}", f"{z[r]:.2f}"][: {11: 1, 111: 2, 1111: 3}[cols]]
            f.write(sep.join([str(2000 + r)] + vals) + ("\n" if r < n - 1 else ""))  # no final newline
    a = su.read_nav(p, cols, dirb, 2500.0)
    b = S.read_nav(p, cols, dirb, 2500.0)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    np.testing.assert_array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    assert a[3:] == b[3:]
    # shot selection by ID (shot_i / shot_f of the parfile)
    c = su.read_nav(p, cols, dirb, 2500.0, shot_i=2010, shot_f=2200)
    assert c[3:] == (2010, 2200, 11, 201)


def test_navigation_rejects_malformed_files(su, tmp_path):
    p = os.path.join(str(tmp_path), "bad.txt")
    with open(p, "w") as f:
        f.write("1 100.0\n2.5 200.0\n")
    with pytest.raises(ValueError):
        su.read_nav(p, 11, 1, 0.0)
    with pytest.raises(ValueError):
        su.read_nav(os.path.join(str(tmp_path), "missing.txt"), 11, 1, 0.0)
    with open(p, "w") as f:
        f.write("1 100.0\n")
    with pytest.raises(ValueError):
        su.read_nav(p, 111, 1, 0.0)  # a second column is missing
