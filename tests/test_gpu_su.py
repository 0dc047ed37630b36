"""GPU parity of the SU codec (psdm_rtm_su_decode / _header_i32 / _encode) with the oracle
(oracle/su_codec.py, pinned against the reference's reader in tests/test_su_oracle.py):
bit for bit, NaN payloads, infinities and denormals included."""
import sys

import numpy as np
import pytest

from oracle import su_codec as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def su_bytes(rng, ntr, ns, big):
    e = ">" if big else "<"
    hdr = rng.integers(-2**31, 2**31 - 1, (ntr, 60)).astype(e + "i4")
    smp = rng.standard_normal((ntr, ns)).astype(np.float32)
    if ns >= 8:
        smp[0, :8] = np.array([0x80000000, 0x7F800000, 0xFF800000, 0x00000001, 0x7FC00001, 0xFFA5A5A5, 0x7F7FFFFF,
                               0x00800000], np.uint32).view(np.float32)
    return b"".join(hdr[j].tobytes() + smp[j].astype(e + "f4").tobytes() for j in range(ntr))


def to_dev(b):
    import torch

    return torch.from_numpy(np.frombuffer(b, np.uint8).copy()).to("cuda:0")


@pytest.mark.parametrize("ntr,ns,nt,big,shift", [
    (480, 2401, 2401, True, 0),    # one MCS gather, 4.8 s at 2 ms
    (480, 2401, 1990, False, 0),   # nt < ns
    (1, 1, 1, True, 0), (3, 1023, 1023, True, 0), (3, 1024, 1024, False, 0), (5, 1025, 1025, True, 0),
    (17, 3000, 2999, True, 37), (17, 3000, 3000, True, -250), (4, 100, 90, False, 90), (4, 100, 90, True, -200)])
def test_decode_bit_exact(ctx, ntr, ns, nt, big, shift):
    from psdm_amd import su

    say(f"[gpu] su decode ntr={ntr} ns={ns} nt={nt} big={big} shift={shift}")
    rng = np.random.default_rng(ntr + ns + nt)
    b = su_bytes(rng, ntr, ns, big)
    out = su.decode(ctx, to_dev(b), ntr, ns, nt, big_endian=big, shift=shift).cpu().numpy()
    np.testing.assert_array_equal(out.view(np.uint32), S.decode(b, ntr, ns, nt, big, shift).view(np.uint32))


def test_decode_whole_line_in_one_call(ctx):
    """64 gathers of 480 traces back to back (the line's raw records) decode in one launch
    to the same bits as gather by gather."""
    import torch

    from psdm_amd import su

    say("[gpu] su decode line 64 x 480 x 2401")
    rng = np.random.default_rng(1)
    ns = 2401
    shots = [su_bytes(rng, 480, ns, True) for _ in range(4)]
    b = b"".join(shots * 16)
    d = to_dev(b)
    line = su.decode(ctx, d, 64 * 480, ns)
    torch.cuda.synchronize()
    for k in (0, 5, 63):
        want = S.decode(shots[k % 4], 480, ns, ns)
        np.testing.assert_array_equal(line[k * 480:(k + 1) * 480].cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("big", [True, False])
def test_header_words(ctx, big):
    from psdm_amd import su

    say(f"[gpu] su header words big={big}")
    rng = np.random.default_rng(7)
    ntr, ns = 700, 333
    b = su_bytes(rng, ntr, ns, big)
    d = to_dev(b)
    for pos in (1, 9, 37, 71, 115, 237, (60 + ns) * 4 - 3):
        got = su.header_i32(ctx, d, ntr, ns, pos, big_endian=big).cpu().numpy()
        np.testing.assert_array_equal(got, S.header_i32(b, ntr, ns, pos, big))


@pytest.mark.parametrize("ni,nj,di,dj,trid,big", [
    (2001, 601, 10.0, 10.0, 1, True),   # the final image of the 20 km line
    (601, 2001, 12.5, 0.004, 130, False), (1, 1, 0.1, 1.0, -5, True), (3, 964, 2.5, 2.5, 1, True),
    (5, 965, 7.0, 1.0, 1, False), (2, 65535, 1.0, 1.0, 1, True)])
def test_encode_bit_exact_and_round_trip(ctx, ni, nj, di, dj, trid, big):
    import torch

    from psdm_amd import su

    say(f"[gpu] su encode ni={ni} nj={nj} big={big}")
    rng = np.random.default_rng(ni + nj)
    img = rng.standard_normal((ni, nj)).astype(np.float32)
    d_img = torch.from_numpy(img).to("cuda:0")
    d_su = su.encode(ctx, d_img, di, dj, trid=trid, big_endian=big)
    got = d_su.cpu().numpy().tobytes()
    assert got == S.encode(img, di, dj, trid, big)
    back = su.decode(ctx, d_su, ni, nj, big_endian=big).cpu().numpy()
    np.testing.assert_array_equal(back.view(np.uint32), img.view(np.uint32))


def test_file_round_trip(ctx, tmp_path):
    """Image -> .su file on disk -> bytes back to the device -> gather."""
    import torch

    from psdm_amd import su

    say("[gpu] su file round trip")
    img = np.random.default_rng(2).standard_normal((50, 120)).astype(np.float32)
    p = str(tmp_path / "img.su")
    su.write_su(p, su.encode(ctx, torch.from_numpy(img).to("cuda:0"), 10.0, 10.0))
    d = su.load_bytes(p)
    np.testing.assert_array_equal(su.decode(ctx, d, 50, 120).cpu().numpy(), img)
    assert su.header_i32(ctx, d, 50, 120, 1).cpu().numpy().tolist() == list(range(1, 51))


def test_rejects_bad_arguments(ctx):
    import torch

    from psdm_amd import su
    from psdm_amd._native import RtmError

    d = torch.zeros(10 * (240 + 400), dtype=torch.uint8, device="cuda:0")
    with pytest.raises(RtmError, match="nbytes"):
        su.decode(ctx, d, 11, 100)
    with pytest.raises(RtmError, match="nt <= ns"):
        su.decode(ctx, d, 10, 100, nt=101)
    with pytest.raises(RtmError, match="aligned"):
        su.decode(ctx, d[1:], 1, 100)
    with pytest.raises(RtmError, match="byte position"):
        su.header_i32(ctx, d, 10, 100, 638)
    # a trace of 240 + 4*ns bytes must stay below 2^31 (int word and byte offsets)
    with pytest.raises(RtmError, match="2\\^29"):
        su.decode(ctx, d, 1, 1 << 29, nt=1)
    with pytest.raises(RtmError, match="2\\^29"):
        su.header_i32(ctx, d, 1, 1 << 29, 9)
    with pytest.raises(RtmError, match="65535"):
        su.encode(ctx, torch.zeros((1, 65536), device="cuda:0"), 1.0, 1.0)
