"""The rocprofv3 CSV summariser used for the kernel-trace and counter figures."""
import csv

from psdm_amd.trace_stats import _short, main, union_ns


def test_union_of_intervals():
    assert union_ns([]) == 0
    assert union_ns([(0, 10), (5, 20), (30, 40)]) == 30
    assert union_ns([(10, 12), (0, 10), (2, 3)]) == 12


def test_trace_busy_time_per_step_of_concurrent_dispatches(tmp_path, capsys):
    p = tmp_path / "kt_kernel_trace.csv"
    rows = [("void psdm::rtm_batch_kernel<16>(a)", 0, 100), ("void psdm::rtm_batch_kernel<16>(b)", 50, 160),
            ("void psdm::rtm_batch_kernel<16>(a)", 200, 280), ("void psdm::rtm_batch_kernel<16>(b)", 210, 300),
            ("psdm::rtm_advance_kernel(x)", 300, 310)]
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writerows(rows)
    main(["trace", str(p), "--kernel", "rtm_batch_kernel", "--chains", "2"])
    out = capsys.readouterr().out
    assert "4 dispatches = 2 steps" in out and "busy time per step 130 ns" in out
    assert "| `psdm::rtm_advance_kernel` | 1 | 10 |" in out


def test_kernel_names_without_arguments():
    assert _short("void psdm::rtm_batch_kernel<16, 0>(psdm::BatchArgs, x)") == "psdm::rtm_batch_kernel<16, 0>"
    assert _short("psdm::(anonymous namespace)::cond_toeplitz_kernel(float const*, int)") == \
        "psdm::cond_toeplitz_kernel"
