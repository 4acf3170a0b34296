"""Algorithmic FLOP model of the ERI path, per canonical class (la lb | lc ld).

The counting rules are the textbook Head-Gordon-Pople scheme, independent of
how the kernels are split or scheduled:

  * set-up: 30 flops per primitive quartet (P, Q, W, rho, T, prefactor);
  * Boys function: 12 flops (Taylor step) + 4 per downward-recursion order;
  * Obara-Saika vertical recurrence [e0|f0]^(m), e in [la, la+lb],
    f in [lc, lc+ld] (and every lower node the recursion needs): 2 flops per
    multiply-add term present; the (x - rho/zeta y) term counts as two;
  * horizontal recurrence: 2 flops per generated element;
  * direct J/K: 6 multiply-adds (12 flops) per integral;
  * adjoint gradient: the derivative class's recurrences (all integrals with
    a, b or c raised or lowered by one in each direction, d/dD from
    translational invariance), the whole class in one DAG; per derivative
    integral (9 per output: A, B, C x xyz) 2 flops to form it from the
    shifted integrals and 2 flops to contract it with the density pair
    product, plus 5 flops per output for that product.

Cartesian-to-spherical transforms act on n x n matrices (the density and
J, K), not on integrals, so they add nothing per quartet.  The DAGs are the
generator's own (gen_eri.build_dag), counted whole: the kernels' executed
counts exceed these where a class is split into parts that recompute shared
recurrence nodes.

Usage: python -m quax_amd.csrc.flop_model [OUT.json]   (default: flop_model.json
beside this file)
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gen_eri as G  # noqa: E402


def vrr_flops(v):
    """2 flops per multiply-add term of every VRR node in v.memo."""
    total = 0
    for (e, f, m), name in v.memo.items():
        if name.startswith("b"):
            continue
        if G.tot(f) > 0:
            i = G.first_axis(f)
            f1 = G.inc(f, i, -1)
            terms = 2 + (2 if f1[i] > 0 else 0) + (1 if e[i] > 0 else 0)
        else:
            i = G.first_axis(e)
            e1 = G.inc(e, i, -1)
            terms = 2 + (2 if e1[i] > 0 else 0)
        total += 2 * terms
    return total


def class_model(l):
    la, lb, lc, ld = l
    L = sum(l)
    N = G.ncart(la) * G.ncart(lb) * G.ncart(lc) * G.ncart(ld)
    T = [(a, b, c, d) for a in G.comps(la) for b in G.comps(lb) for c in G.comps(lc) for d in G.comps(ld)]
    h, v, _, _, _ = G.build_dag(T)
    eri = 30 + 12 + 4 * L + vrr_flops(v) + 2 * len(h.lines)
    whole = {"combos": [(c, x) for c in "ABC" for x in range(3)], "dinv": True, "a": (0, G.ncart(la)),
             "b": (0, G.ncart(lb)), "c": (0, G.ncart(lc))}
    hd, vd, _, _, _ = G.build_dag(G.deriv_targets(l, whole))
    deri = 30 + 12 + 4 * (L + 1) + vrr_flops(vd) + 2 * len(hd.lines) + N * (9 * (2 + 2) + 5)
    return {"N": N, "eri": eri, "jk": eri + 12 * N, "adjoint": deri}


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "flop_model.json")
    _, classes = G.canonical_classes(2)
    model = {"".join(map(str, c)): class_model(c) for c in classes}
    with open(out, "w") as fh:
        json.dump({"rules": __doc__.split("Usage:")[0].strip(), "classes": model}, fh, indent=1)
    print(out)


if __name__ == "__main__":
    main()
