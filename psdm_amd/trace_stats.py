"""Summaries of rocprofv3 CSV output for the RTM kernels.

    python -m psdm_amd.trace_stats trace KERNEL_TRACE.csv [--kernel rtm_batch_kernel --chains 2]
    python -m psdm_amd.trace_stats pmc COUNTER_COLLECTION.csv [--kernel rtm_batch_kernel --chains 2]

trace: dispatch count and mean / min / max duration per kernel.  With --kernel, also the
busy time per step: the union of that kernel's dispatch intervals divided by the number
of steps (dispatches / chains).  The engine issues one step of its resident shots as
`chains` concurrent dispatches (one per slot group), so a step lasts from the first of
them starting to the last ending, and the union is the sum of those spans.

pmc: per kernel and counter, the mean value per dispatch; with --kernel, also the value
per step (sum over the step's `chains` dispatches).
"""
from __future__ import annotations

import argparse
import csv
import sys
from collections import defaultdict


def _short(name: str) -> str:
    """Kernel name without its argument list ("(anonymous namespace)::" dropped first)."""
    return name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")


def union_ns(intervals) -> int:
    """Total length of the union of [start, end) intervals."""
    tot, cur_s, cur_e = 0, None, None
    for s, e in sorted(intervals):
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                tot += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    if cur_e is not None:
        tot += cur_e - cur_s
    return tot


def trace(path: str, kernel: str | None, chains: int) -> None:
    per = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            per[_short(r["Kernel_Name"])].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    total = sum(e - s for v in per.values() for s, e in v)
    print("| kernel | calls | average ns | min ns | max ns | share |")
    print("|---|---|---|---|---|---|")
    for nm, v in sorted(per.items(), key=lambda kv: -sum(e - s for s, e in kv[1])):
        d = [e - s for s, e in v]
        print(f"| `{nm}` | {len(d)} | {sum(d) / len(d):.0f} | {min(d)} | {max(d)} | {100 * sum(d) / total:.2f} % |")
    if kernel:
        v = [iv for nm, ivs in per.items() if kernel in nm for iv in ivs]
        if not v:
            sys.exit(f"no dispatch of {kernel!r} in {path}")
        steps = len(v) / chains
        print(f"\n{kernel}: {len(v)} dispatches = {steps:.0f} steps of {chains} concurrent dispatches; "
              f"busy time per step {union_ns(v) / steps:.0f} ns (mean dispatch {sum(e - s for s, e in v) / len(v):.0f} ns)")


def pmc(path: str, kernel: str | None, chains: int) -> None:
    per = defaultdict(lambda: defaultdict(list))
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            per[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    for nm, cs in sorted(per.items()):
        for cn, v in sorted(cs.items()):
            line = f"{nm} {cn}: {len(v)} dispatches, mean {sum(v) / len(v):.1f} per dispatch"
            if kernel and kernel in nm:
                line += f", {chains * sum(v) / len(v):.1f} per step ({chains} dispatches)"
            print(line)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--chains", type=int, default=1)
    a = ap.parse_args(argv)
    (trace if a.what == "trace" else pmc)(a.csv, a.kernel, a.chains)


if __name__ == "__main__":
    main()
