"""Writes tests/golden/conditioning_ref.npz: gathers conditioned by the reference's own
routines (oracle/_ref: freq_filter with FFTPACK, interpolation_akima, taper_pad_trace)
in the driver's order (calculate_PSDM_RTM.f90:299-358).  Run in the container where the
reference is mounted and oracle/_ref is built (make -C oracle ref):

    python -m tests.golden.make_conditioning_golden
"""
import os

import numpy as np

from oracle import conditioning as O

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = ["dt_data", "typef", "f_init", "f_final", "nt_prop2", "dt_prop", "nt_taper", "N_taper", "N_zeros",
          "taper_option"]
CASES = {
    "bandpass_resample_taper": dict(L=500, n=8, dt_data=0.002, typef=3, f_init=5.0, f_final=40.0, nt_prop2=1200,
                                    dt_prop=0.0008, nt_taper=1250, N_taper=40, N_zeros=50, taper_option=1, sf=True),
    "lowpass_cut_multiple": dict(L=400, n=6, dt_data=0.004, typef=1, f_init=2.0, f_final=25.0, nt_prop2=1000,
                                 dt_prop=0.0015, nt_taper=1030, N_taper=25, N_zeros=30, taper_option=0, cut=True),
    "highpass_same_step": dict(L=500, n=5, dt_data=0.002, typef=2, f_init=5.0, f_final=60.0, nt_prop2=450,
                               dt_prop=0.002, nt_taper=480, N_taper=16, N_zeros=30, taper_option=1),
    "resample_only_down": dict(L=900, n=4, dt_data=0.0008, typef=3, f_init=0.0, f_final=0.0, nt_prop2=330,
                               dt_prop=0.002, nt_taper=330, N_taper=1, N_zeros=0, taper_option=0, sf=True),
}


def reference(raw, c, nt_sf, nt_cut):
    d = O.ref_freq_filter(raw, c["dt_data"], c["typef"], c["f_init"], c["f_final"]) if c["f_init"] > 0 else raw
    n = raw.shape[1]
    out = np.zeros((c["nt_taper"], n), np.float32, order="F")
    if np.float32(c["dt_data"]) != np.float32(c["dt_prop"]):
        out[:c["nt_prop2"]] = O.ref_akima(d, c["dt_data"], c["nt_prop2"], c["dt_prop"])
    else:
        out[:c["nt_prop2"]] = d[:c["nt_prop2"]]
    for i in range(n):
        if nt_sf is not None:
            out[:nt_sf[i], i] = 0
        if nt_cut is not None:
            out[nt_cut[i] - 1:, i] = 0
            out[:, i] = O.ref_taper_pad_trace(out[:, i], int(nt_cut[i]), c["N_taper"], c["N_zeros"])
        elif c["taper_option"]:
            out[:, i] = O.ref_taper_pad_trace(out[:, i], c["nt_prop2"], c["N_taper"], c["N_zeros"])
    return out


def main():
    rng = np.random.default_rng(2024)
    arrays = {"param_names": np.array(PARAMS)}
    for name, c in CASES.items():
        t = np.arange(c["L"]) * c["dt_data"]
        raw = np.zeros((c["L"], c["n"]), np.float32, order="F")
        for i in range(c["n"]):  # a few Ricker-like arrivals plus noise
            for t0 in rng.uniform(0.05, 0.8 * t[-1], 3):
                a = (np.pi * 20.0 * (t - t0)) ** 2
                raw[:, i] += rng.uniform(-1, 1) * (1 - 2 * a) * np.exp(-a)
            raw[:, i] += 0.05 * rng.standard_normal(c["L"])
        raw[:, -1] = 0.0  # an all-zero trace
        nt_sf = rng.integers(0, c["nt_prop2"] // 3, c["n"]).astype(np.int32) if c.get("sf") else None
        nt_cut = rng.integers(c["nt_prop2"] // 4, c["nt_taper"] + 1, c["n"]).astype(np.int32) if c.get("cut") else None
        arrays[name + "_raw"] = raw
        arrays[name + "_params"] = np.array([c[k] for k in PARAMS], np.float64)
        if nt_sf is not None:
            arrays[name + "_nt_sf"] = nt_sf
        if nt_cut is not None:
            arrays[name + "_nt_cut"] = nt_cut
        arrays[name + "_out"] = reference(raw, c, nt_sf, nt_cut)
    np.savez_compressed(os.path.join(HERE, "conditioning_ref.npz"), **arrays)


if __name__ == "__main__":
    main()
