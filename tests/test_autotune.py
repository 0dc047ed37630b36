"""The autotuner's round model (psdm_amd/autotune.py) on the CPU: trials replaced by a
cost model in which a step costs a fixed launch overhead plus a part per busy slot, so a
round with few busy slots costs more per shot than a full one."""
import pytest

from psdm_amd import autotune as A


def test_slot_candidates_split_the_shots_into_equal_rounds():
    assert A.slot_candidates(20) == [10, 16, 20, 24, 32]
    assert A.slot_candidates(64) == [16, 22, 24, 32]
    assert A.slot_candidates(None) == [16, 24, 32]


def test_round_time_counts_the_partial_round_at_its_own_cost():
    # 20 shots through 16 slots: one full round (t = 42) and one of 4 (t = 18)
    assert A.round_time(42.0, 20, 16, 18.0) == pytest.approx(3.0)
    assert A.round_time(50.0, 20, 20) == pytest.approx(2.5)
    assert A.round_time(50.0, 40, 20) == pytest.approx(2.5)


def test_tune_prefers_full_rounds(monkeypatch):
    calls = []

    def fake_trial(ctx, d_vel, ny, nx_prop, jobs, nslots, busy, max_nx, nt, dPML, fs, dx, dt, cfg, *rest):
        calls.append((nslots, busy, cfg["tile_x"], cfg["stream_w"]))
        if nslots == 32:
            raise RuntimeError("out of memory")
        shape_cost = 0.0 if (cfg["tile_x"], cfg["stream_w"]) == (16, 4) else 5e-6
        return 10e-6 + 2e-6 * busy + 0.01e-6 * nslots + shape_cost

    monkeypatch.setattr(A, "_trial", fake_trial)
    best, trials = A.tune(None, None, 601, 2001, [{}], 715, 6000, 40, 0, 10.0, 8e-4, nshots=20)
    assert (best["tile_x"], best["stream_w"]) == (16, 4)
    # 16 slots: (42.16 + 18.16) / 20 = 3.016 us per shot-step; 20 slots: 50.2 / 20 = 2.51
    assert best["nslots"] == 20
    assert best["score"] == pytest.approx(2.51)
    assert (16, 4, 16, 4) in calls  # the partial round of 16 slots was timed at 4 busy
    assert not any(t.get("nslots") == 32 for t in trials)


def test_tune_plain_picks_the_fastest_shape_and_leaves_it_set():
    """tune_plain on a stand-in context whose step kernel time depends on the launch shape:
    every (tile, width) is probed twice (once untimed), the fastest is chosen, and it is the
    shape left on the context."""

    class Ctx:
        def __init__(self):
            self.shape = [8, 2]
            self.timing = 0
            self.runs = []

        def set_launch(self, tile_x=8):
            self.shape[0] = tile_x

        def set_stream_width(self, w):
            self.shape[1] = w

        def set_timing(self, n):
            self.timing = n

        def sync(self):
            pass

        def step_stats(self):
            tx, sw = self.shape
            return 100, 0.01 + 0.001 * abs(tx - 16) / 8 + 0.002 * abs(sw - 2)

    ctx = Ctx()

    def probe():
        ctx.runs.append((tuple(ctx.shape), ctx.timing))

    best, trials = A.tune_plain(ctx, probe)
    assert (best["tile_x"], best["stream_w"]) == (16, 2)
    assert ctx.shape == [16, 2] and ctx.timing == 0
    assert len(trials) == 12 and len(ctx.runs) == 24
    assert all(((t["tile_x"], t["stream_w"]), 0) in ctx.runs for t in trials)
