"""Derivative-part autotuner: picks, per angular-momentum class, the node
budget at which gen_eri.py splits the adjoint-gradient kernel (--key budget)
or the waves per SIMD its kernels are compiled for (--key waves; a library
built with --waves W, W = 0 for the compiler's choice).

A smaller budget splits a class into more parts: each part recomputes the
recurrence nodes it shares with the others (more FP64 work) but keeps its
live set in registers (less scratch traffic).  Which side wins depends on the
class, so it is measured.  Workflow:

  1. on a build host, one library per candidate budget:
       python -m quax_amd.csrc.build --no-fp32 --budget B --out DIR/B/libquax_hip.so
  2. on the GPU:
       python -m quax_amd.csrc.tune --out deriv_budget.json DIR/B1/libquax_hip.so ...
     runs bench.py on one stream with each library (QUAX_AMD_LIB) and keeps,
     per class, the budget with the lowest summed adjoint-kernel time;
  3. copy the JSON next to gen_eri.py (quax_amd/csrc/deriv_budget.json) and
     rebuild: gen_eri.py reads it.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def budget_of(lib_path: str) -> int:
    """The tuned value a library was built with: the number in its directory name."""
    m = re.search(r"(\d+)", os.path.basename(os.path.dirname(os.path.abspath(lib_path))))
    if not m:
        raise SystemExit(f"cannot read the budget from the directory name of {lib_path}")
    return int(m.group(1))


def measure(lib_path: str, workload: str, steps: int, timeout: int) -> dict:
    env = dict(os.environ, QUAX_AMD_LIB=os.path.abspath(lib_path))
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--workload", workload, "--steps", str(steps),
           "--warmup", "1", "--cpu-seconds", "0", "--no-scf", "--no-mixed", "--streams", "1"]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise SystemExit(f"bench.py failed with {lib_path}:\n{r.stderr[-3000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("libs", nargs="+", help="one library per budget, in directories named by the budget")
    ap.add_argument("--workload", default="cluster10")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--timeout", type=int, default=300)
    ap.add_argument("--out", default="deriv_budget.json")
    ap.add_argument("--key", default="budget", choices=["budget", "waves"])
    ap.add_argument("--merge", default=None, help="existing tuning JSON whose other keys are kept")
    a = ap.parse_args()
    per = {}
    for lib in a.libs:
        b = budget_of(lib)
        res = measure(lib, a.workload, a.steps, a.timeout)
        per[b] = {cls: t.get("deriv64", 0.0) for cls, t in res["class_kernel_ms_per_step"].items()}
        print(f"[tune] {a.key} {b}: adjoint kernels {sum(per[b].values()):.2f} ms per step", flush=True)
    classes = sorted(set().union(*[set(v) for v in per.values()]))
    choice, table = {}, {}
    for cls in classes:
        times = {b: per[b][cls] for b in per if cls in per[b]}
        best = min(times, key=times.get)
        choice[cls] = best
        table[cls] = times
        print(f"[tune] {cls}: " + "  ".join(f"{b}:{times[b]:.3f}" for b in sorted(times)) + f"  -> {best}",
              flush=True)
    out = {}
    if a.merge and os.path.exists(a.merge):
        with open(a.merge) as fh:
            out = json.load(fh)
    out[a.key] = choice
    out[f"ms_per_step_{a.key}"] = table
    out["workload"] = a.workload
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"[tune] wrote {a.out}")


if __name__ == "__main__":
    main()
