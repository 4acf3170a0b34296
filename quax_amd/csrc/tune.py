"""Autotuner of the per-class kernel choices, from GPU timings.

--tiles (in process, one library): the quartet tile of every class -- the
threads per workgroup of its streaming J/K and adjoint-gradient kernels,
i.e. how many primitive quartets a workgroup evaluates and how many share its
LDS K' rows -- out of 64/128/192/256, timed per class on one stream, and
the classes whose adjoint-gradient kernels gain from summing wave-uniform
terms across the wavefront, and the classes whose FP32 kernels beat their
FP64 ones by 8% or more (the mixed-precision policy).  Writes the "tiles",
"grad_wave_reduce" and "fp32" keys that gen_eri.py compiles into the default
tables (qx_set_class_tiles / qx_set_grad_wave_reduce / qx_set_fp32_classes
override them at run time):
    python -m quax_amd.csrc.tune --tiles --merge quax_amd/csrc/deriv_budget.json --out tiles.json

--jk-rows (in process): per class, the J/K kernels' time with K' rows staged
in LDS against every K' term added to global memory (a measurement only):
    python -m quax_amd.csrc.tune --jk-rows

Otherwise (several libraries): the node budget at which gen_eri.py splits
the adjoint-gradient kernel (--key budget) or the waves per SIMD its kernels
are compiled for (--key waves; a library built with --waves W, W = 0 for the
compiler's choice).

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


TILE_CHOICES = (64, 128, 192, 256)
FP32_GAIN = 0.92   # FP32 admitted when its kernels take at most this fraction of the FP64 time


def _class_names(L):
    import ctypes
    names = []
    for c in range(L.qx_num_classes()):
        l4 = (ctypes.c_int * 4)()
        npart = ctypes.c_int()
        L.qx_class_info(c, l4, ctypes.byref(npart))
        names.append("".join(str(x) for x in l4))
    return names


def _pass_times(eng, D, steps, precision=0, thr=0.0):
    """Kernel ms per step per (class name, pass) from HIP-event profiling on
    one stream; pass 0 = J/K kernels, 1 = adjoint-gradient kernels."""
    import ctypes
    import torch
    L = eng.L
    names = _class_names(L)
    J, K = torch.empty_like(D), torch.empty_like(D)

    def step():
        eng.jk(D, precision=precision, fp32_threshold=thr, out=(J, K))
        eng.eri_grad(D, precision=precision, fp32_threshold=thr)

    step()
    torch.cuda.synchronize()
    eng.profile(True)
    for _ in range(steps):
        step()
    ms, _, _ = eng.profile_read()
    eng.profile(False)
    out = {}
    for slot, t in enumerate(ms):
        if t <= 0:
            continue
        kind, cls, part, f32 = (ctypes.c_int() for _ in range(4))
        L.qx_profile_slot(slot, ctypes.byref(kind), ctypes.byref(cls), ctypes.byref(part), ctypes.byref(f32))
        if kind.value not in (1, 2):
            continue
        key = (names[cls.value], 0 if kind.value == 2 else 1)
        out[key] = out.get(key, 0.0) + float(t) / steps
    return out


def tune_tiles(workload, steps, seed=0):
    import numpy as np
    import torch
    from quax_amd import workloads
    from quax_amd.basis import BasisSet
    from quax_amd.engine import EriEngine
    mol = getattr(workloads, workload)()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    eng = EriEngine(mol, bs, ints_tol=1e-12)
    n = eng.nbf
    X = np.random.default_rng(seed).standard_normal((n, n)) * 0.3
    D = torch.as_tensor(X @ X.T / n, device=eng.device)
    nc = eng.L.qx_num_classes()
    names = _class_names(eng.L)
    eng.set_streams(1)
    per = {}
    for b in TILE_CHOICES:
        for p in (0, 1):
            eng.set_class_tiles(p, [b] * nc)
        per[b] = _pass_times(eng, D, steps)
        print(f"[tune] tile {b}: J/K {sum(v for (c, p), v in per[b].items() if p == 0):.2f} ms, adjoint "
              f"{sum(v for (c, p), v in per[b].items() if p == 1):.2f} ms per step (one stream)", flush=True)
    tiles = {"jk": {}, "grad": {}}
    table = {}
    for (cls, p) in sorted(set().union(*[set(v) for v in per.values()])):
        times = {b: per[b][(cls, p)] for b in per if (cls, p) in per[b]}
        best = min(times, key=times.get)
        # keep the default unless another tile is at least 2% faster
        if times.get(128, 1e300) <= 1.02 * times[best]:
            best = 128
        tiles["jk" if p == 0 else "grad"][cls] = best
        table[f"{cls}/{'jk' if p == 0 else 'grad'}"] = times
        print(f"[tune] {cls} {'J/K ' if p == 0 else 'grad'}: " + "  ".join(f"{b}:{times[b]:.3f}" for b in sorted(times))
              + f"  -> {best}", flush=True)
    for p, key in ((0, "jk"), (1, "grad")):
        eng.set_class_tiles(p, [tiles[key].get(names[c], 128) for c in range(nc)])
    # copies of the gradient accumulators the blocks flush into
    gcp = {}
    for cp in (1, 8, 64):
        eng.set_grad_copies([cp] * nc)
        gcp[cp] = _pass_times(eng, D, steps)
    copies = {}
    for (cls, p) in sorted(gcp[64]):
        if p != 1:
            continue
        times = {cp: gcp[cp].get((cls, p), 1e300) for cp in gcp}
        best = min(times, key=times.get)
        if times[64] <= 1.02 * times[best]:
            best = 64
        copies[cls] = best
        table[f"{cls}/grad_copies"] = times
    eng.set_grad_copies([copies.get(names[c], 64) for c in range(nc)])
    print(f"[tune] gradient copies: {copies}", flush=True)
    # wave reduction of the gradient terms: every class with and without it
    wred = {}
    for on in (False, True):
        eng.set_grad_wave_reduce([on] * nc)
        wred[on] = _pass_times(eng, D, steps)
    gwr = sorted(cls for (cls, p), t in wred[True].items()
                 if p == 1 and t <= 0.98 * wred[False].get((cls, p), 1e300))
    eng.set_grad_wave_reduce([names[c] in gwr for c in range(nc)])
    print(f"[tune] wave-reduced gradient terms: {gwr}", flush=True)
    for (cls, p) in sorted(wred[True]):
        if p == 1:
            table[f"{cls}/grad_wave_reduce"] = {"off": wred[False].get((cls, p)), "on": wred[True][(cls, p)]}
    # FP32 policy: every quartet of every class in FP32 against FP64, at the chosen tiles
    t64 = _pass_times(eng, D, steps)
    for p in (0, 1):
        eng.set_fp32_classes(p, np.ones(nc, dtype=bool))
    t32 = _pass_times(eng, D, steps, precision=1, thr=1e300)
    fp32 = {"jk": [], "grad": []}
    for (cls, p), t in sorted(t64.items()):
        if (cls, p) in t32 and t32[(cls, p)] <= FP32_GAIN * t:
            fp32["jk" if p == 0 else "grad"].append(cls)
    print(f"[tune] FP32 admitted: J/K {fp32['jk']}, gradient {fp32['grad']}", flush=True)
    # the whole step on four streams: tuned tiles against the 128 default
    eng.set_streams(4)
    for p in (0, 1):
        eng.set_fp32_classes(p, None)
    res = {}
    for label, tl in (("default", None), ("tuned", tiles)):
        for p, key in ((0, "jk"), (1, "grad")):
            eng.set_class_tiles(p, [128] * nc if tl is None else [tl[key].get(names[c], 128) for c in range(nc)])
        J, K = torch.empty_like(D), torch.empty_like(D)
        for _ in range(2):
            eng.jk(D, out=(J, K))
            eng.eri_grad(D)
        torch.cuda.synchronize()
        import time
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.jk(D, out=(J, K))
            eng.eri_grad(D)
        torch.cuda.synchronize()
        res[label] = (time.perf_counter() - t0) / steps * 1e3
        print(f"[tune] four streams, {label} tiles: {res[label]:.2f} ms per step", flush=True)
    eng.close()
    return {"tiles": {k: {c: b for c, b in v.items() if b != 128} for k, v in tiles.items()}, "fp32": fp32,
            "grad_wave_reduce": gwr, "grad_copies": {c: n for c, n in copies.items() if n != 64},
            "ms_per_step_tiles": {"per_class_one_stream": table, "four_streams": res}}


def compare_jk_rows(workload, steps, seed=0):
    """Per class, J/K kernel ms per pass on one stream with the K' rows of a
    block's bra pair staged in LDS (qx_set_jk_rows on) and with every K' term
    added to global memory (off); best of two rounds each."""
    import numpy as np
    import torch
    from quax_amd import workloads
    from quax_amd.basis import BasisSet
    from quax_amd.engine import EriEngine
    mol = getattr(workloads, workload)()
    bs = BasisSet.build(mol, "even-tempered", pure=True)
    with EriEngine(mol, bs, ints_tol=1e-12) as eng:
        eng.set_streams(1)
        n = eng.nbf
        X = np.random.default_rng(seed).standard_normal((n, n)) * 0.3
        D = torch.as_tensor(X @ X.T / n, device=eng.device)
        nc = eng.L.qx_num_classes()
        names = _class_names(eng.L)
        res = {}
        for _ in range(2):
            for on in (1, 0):
                eng.set_jk_rows([on] * nc)
                t = _pass_times(eng, D, steps)
                for nm in names:
                    res.setdefault((nm, on), []).append(t.get((nm, 0), 0.0))
        eng.set_jk_rows(None)
        ton = toff = 0.0
        for nm in names:
            a, b = min(res[(nm, 1)]), min(res[(nm, 0)])
            ton += a
            toff += b
            print(f"[tune] {nm} J/K: K' rows in LDS {a:.3f} ms, direct {b:.3f} ms ({b / a:.2f}x)", flush=True)
        print(f"[tune] J/K pass: K' rows in LDS {ton:.2f} ms, direct {toff:.2f} ms", flush=True)


def main():
    if "--jk-rows" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--jk-rows", action="store_true")
        ap.add_argument("--workload", default="cluster10")
        ap.add_argument("--steps", type=int, default=4)
        a = ap.parse_args()
        compare_jk_rows(a.workload, a.steps)
        return
    if "--tiles" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--tiles", action="store_true")
        ap.add_argument("--workload", default="cluster10")
        ap.add_argument("--steps", type=int, default=3)
        ap.add_argument("--out", default="tiles.json")
        ap.add_argument("--merge", default=None, help="existing tuning JSON whose other keys are kept")
        a = ap.parse_args()
        res = tune_tiles(a.workload, a.steps)
        out = {}
        if a.merge and os.path.exists(a.merge):
            with open(a.merge) as fh:
                out = json.load(fh)
        fp_note = out.get("fp32", {}).get("note")
        out.update(res)
        out["fp32"]["note"] = (f"classes whose FP32 kernels took <= {FP32_GAIN:.2f} of the FP64 time "
                               f"({a.workload}, one stream, whole lists, tuned tiles)")
        del fp_note
        out["tiles_workload"] = a.workload
        with open(a.out, "w") as fh:
            json.dump(out, fh, indent=1, sort_keys=True)
        print(f"[tune] wrote {a.out}")
        return
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
