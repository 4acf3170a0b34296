"""Generates the per-angular-momentum-class ERI and ERI-derivative kernels.

One ERI kernel per canonical class (la lb | lc ld) with la >= lb, lc >= ld and
bra pair class >= ket pair class.  Each thread evaluates one primitive-pair
quartet (contracted shells are handled by summing primitive-pair quartets in
the epilogue), as straight-line code:

  1. setup (P, Q, W, rho) and the Boys function F_0..F_L,
  2. Obara-Saika vertical recurrence [e0|f0]^(m), emitted as SSA scalars in
     dependency order (bra side first, then transferred to the ket),
  3. Head-Gordon-Pople horizontal recurrence to (ab|cd), choosing the cheaper
     of bra-first / ket-first per class,
  4. an epilogue selected at run time (store, full-tensor scatter, Schwarz
     diagonal, direct J/K).

Derivative kernels build d/dX_i (ab|cd) from d/dA_i phi_a = 2 alpha
phi_{a+1_i} - a_i phi_{a-1_i}.  Every kernel forms d/dA, d/dB, d/dC and
infers d/dD by translational invariance.  A class whose body would exceed
its node budget is split into parts over the Cartesian components of a
(then of b, then of c): each part evaluates only the output functions
(a, b, c, d) in its ranges.  A part recomputes the recurrence
nodes it shares with the others, but its live set fits the register file
(the unsplit heavy bodies spill several KB per lane to scratch, and scratch
latency, not arithmetic, then bounds them).  Epilogues: full-tensor scatter
or the adjoint gradient.

Usage: python gen_eri.py OUTDIR [LMAX] [--no-fp32]
"""
from __future__ import annotations

import json
import os
import sys

AX = "xyz"
# largest SSA node count allowed in one derivative body before it is split:
# a per-class table (picked by timing budget variants on the GPU), else
# DERIV_NODES; QX_DERIV_NODES in the environment overrides both (tuning builds)
DERIV_NODES = 2000
# the tuning tables (QX_TUNING_FILE names another file: tests, tuning builds)
_BUDGET_FILE = os.environ.get("QX_TUNING_FILE") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "deriv_budget.json")
DERIV_BUDGET = {}
if os.path.exists(_BUDGET_FILE):  # written by tune.py from GPU timings
    with open(_BUDGET_FILE) as _fh:
        DERIV_BUDGET = {tuple(int(ch) for ch in k): int(v) for k, v in json.load(_fh).get("budget", {}).items()}
_BUDGET_OVERRIDE = os.environ.get("QX_DERIV_NODES")
# waves per SIMD the adjoint kernels of a class are compiled for (0: the
# compiler's choice, one wave with 512 registers for the heavy classes);
# from the same tuning file, QX_GRAD_WAVES overrides (tuning builds)
DERIV_WAVES = {}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        DERIV_WAVES = {tuple(int(ch) for ch in k): int(v) for k, v in json.load(_fh).get("waves", {}).items()}
_WAVES_OVERRIDE = os.environ.get("QX_GRAD_WAVES")


DERIV_CSPLIT = {}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        DERIV_CSPLIT = {tuple(int(ch) for ch in k): int(v) for k, v in json.load(_fh).get("csplit", {}).items()}
_CSPLIT_OVERRIDE = os.environ.get("QX_CSPLIT")

# threads per block of the streaming J/K and gradient kernels per class
# (quartet tiles picked by tune.py --tiles from GPU timings; default 128)
TILES = {"jk": {}, "grad": {}}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        _t = json.load(_fh).get("tiles", {})
        TILES = {k: {tuple(int(ch) for ch in c): int(v) for c, v in _t.get(k, {}).items()} for k in ("jk", "grad")}

# classes whose adjoint-gradient kernels sum wave-uniform centres across the
# wavefront before the LDS accumulators (picked by tune.py --tiles)
GRAD_WAVE_REDUCE = {(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (2, 0, 0, 0)}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        _g = json.load(_fh).get("grad_wave_reduce")
        if _g is not None:
            GRAD_WAVE_REDUCE = {tuple(int(ch) for ch in c) for c in _g}

# copies of the gradient accumulators each class's adjoint blocks flush into
# (picked by tune.py --tiles; default 64)
GRAD_COPIES = {}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        GRAD_COPIES = {tuple(int(ch) for ch in c): int(v)
                       for c, v in json.load(_fh).get("grad_copies", {}).items()}

# classes whose mixed-precision passes may use FP32 (J/K, gradient): those
# whose FP32 kernels timed faster than their FP64 ones (fewer spills)
FP32_CLASSES = {"jk": set(), "grad": set()}
if os.path.exists(_BUDGET_FILE):
    with open(_BUDGET_FILE) as _fh:
        _f = json.load(_fh).get("fp32", {})
        FP32_CLASSES = {k: {tuple(int(ch) for ch in c) for c in _f.get(k, [])} for k in ("jk", "grad")}


def deriv_csplit(l):
    """c-component ranges per (a, b) part for classes whose (a, b) parts
    still exceed their budget (1 = no further split)."""
    if _CSPLIT_OVERRIDE:
        return int(_CSPLIT_OVERRIDE)
    return DERIV_CSPLIT.get(tuple(l), 1)


def deriv_waves(l):
    if _WAVES_OVERRIDE:
        return int(_WAVES_OVERRIDE)
    return DERIV_WAVES.get(tuple(l), 0)


def comps(l):
    out = []
    for i in range(l + 1):
        lx = l - i
        for j in range(i + 1):
            out.append((lx, i - j, j))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


def tot(v):
    return v[0] + v[1] + v[2]


def inc(v, i, d=1):
    w = list(v)
    w[i] += d
    return tuple(w)


def first_axis(v):
    for i in range(3):
        if v[i] > 0:
            return i
    return -1


Z3 = (0, 0, 0)


class VRR:
    """Obara-Saika vertical recurrence DAG for one primitive quartet."""

    def __init__(self):
        self.memo = {}
        self.lines = []
        self.nflop = 0

    def get(self, e, f, m):
        key = (e, f, m)
        if key in self.memo:
            return self.memo[key]
        if tot(f) > 0:
            i = first_axis(f)
            f1 = inc(f, i, -1)
            terms = [f"QC{AX[i]} * {self.get(e, f1, m)}", f"WQ{AX[i]} * {self.get(e, f1, m + 1)}"]
            self.nflop += 3
            if f1[i] > 0:
                f2 = inc(f1, i, -1)
                terms.append(f"R({f1[i]}) * oo2q * ({self.get(e, f2, m)} - rq * {self.get(e, f2, m + 1)})")
                self.nflop += 5
            if e[i] > 0:
                terms.append(f"R({e[i]}) * oo2pq * {self.get(inc(e, i, -1), f1, m + 1)}")
                self.nflop += 3
        elif tot(e) > 0:
            i = first_axis(e)
            e1 = inc(e, i, -1)
            terms = [f"PA{AX[i]} * {self.get(e1, Z3, m)}", f"WP{AX[i]} * {self.get(e1, Z3, m + 1)}"]
            self.nflop += 3
            if e1[i] > 0:
                e2 = inc(e1, i, -1)
                terms.append(f"R({e1[i]}) * oo2p * ({self.get(e2, Z3, m)} - rp * {self.get(e2, Z3, m + 1)})")
                self.nflop += 5
        else:
            name = f"b{m}"
            self.memo[key] = name
            return name
        name = f"v{len(self.memo)}"
        self.memo[key] = name
        self.lines.append(f"const R {name} = {' + '.join(terms)};")
        return name


class HRR:
    """Head-Gordon-Pople horizontal recurrence DAG.  order 'b' reduces b
    before d (so the ket HRR executes first), order 'd' the reverse."""

    def __init__(self, order):
        self.order = order
        self.memo = {}
        self.lines = []
        self.need = set()
        self.nflop = 0

    def get(self, a, b, c, d):
        key = (a, b, c, d)
        if key in self.memo:
            return self.memo[key]
        red_b = tot(b) > 0 and (self.order == "b" or tot(d) == 0)
        if red_b:
            i = first_axis(b)
            b1 = inc(b, i, -1)
            expr = f"{self.get(inc(a, i), b1, c, d)} + AB{AX[i]} * {self.get(a, b1, c, d)}"
            self.nflop += 2
        elif tot(d) > 0:
            i = first_axis(d)
            d1 = inc(d, i, -1)
            expr = f"{self.get(a, b, inc(c, i), d1)} + CD{AX[i]} * {self.get(a, b, c, d1)}"
            self.nflop += 2
        else:
            self.need.add((a, c))
            name = f"s_{a[0]}{a[1]}{a[2]}_{c[0]}{c[1]}{c[2]}"
            self.memo[key] = name
            return name
        name = f"h{len(self.memo)}"
        self.memo[key] = name
        self.lines.append(f"const R {name} = {expr};")
        return name


def build_dag(targets):
    """HRR + VRR DAG for the targets; also returns, per target, the number of
    HRR lines after which its value is available (emission is in target
    order, so consumers can be interleaved to keep live ranges short)."""
    best = None
    for order in "bd":
        h = HRR(order)
        outs, pos = [], []
        for t in targets:
            outs.append(h.get(*t))
            pos.append(len(h.lines))
        v = VRR()
        vn = {ac: v.get(ac[0], ac[1], 0) for ac in sorted(h.need)}
        cost = len(h.lines) + len(v.lines)
        if best is None or cost < best[0]:
            best = (cost, h, v, vn, outs, pos)
    return best[1:]


def interleave(hlines, groups):
    """groups: list of (position, [statements]) in non-decreasing position;
    returns HRR lines with each group's statements placed as soon as its
    inputs exist."""
    out = []
    gi = 0
    for i in range(len(hlines) + 1):
        while gi < len(groups) and groups[gi][0] <= i:
            out += groups[gi][1]
            gi += 1
        if i < len(hlines):
            out.append(hlines[i])
    return out


def emit_dblocks(L, nn, arrays=True):
    """Loads the six density blocks of the quartet (double) from A_.e.D into
    local arrays; arrays=False only sets up the indices (the caller reads D
    where each value is used)."""
    NA, NB, NC, ND = nn
    L.append("  const int n_ = A_.e.nbf;")
    L.append("  const double* Dm_ = A_.e.D;")
    L.append("  const int ia_ = A_.pr.bfa[P1], ib_ = A_.pr.bfb[P1], ic_ = A_.pr.bfa[P2], id_ = A_.pr.bfb[P2];")
    if not arrays:
        return
    for nm, r0, nr, c0, nc in (("Dab", "ia_", NA, "ib_", NB), ("Dcd", "ic_", NC, "id_", ND),
                               ("Dac", "ia_", NA, "ic_", NC), ("Dad", "ia_", NA, "id_", ND),
                               ("Dbc", "ib_", NB, "ic_", NC), ("Dbd", "ib_", NB, "id_", ND)):
        L.append(f"  double {nm}[{nr * nc}];")
        L.append(f"  for (int r_ = 0; r_ < {nr}; ++r_) for (int c_ = 0; c_ < {nc}; ++c_) "
                 f"{nm}[r_ * {nc} + c_] = Dm_[({r0} + r_) * n_ + {c0} + c_];")


CENTRES = "ABCD"


def deriv_targets(l, part):
    """Targets of one derivative part, in emission order: for every
    (a,b,c,d) of the part's a/b ranges, for every (centre, x): raised, then
    lowered."""
    la, lb, lc, ld = l
    a0, a1 = part["a"]
    b0, b1 = part["b"]
    c0, c1 = part.get("c", (0, ncart(lc)))
    T = []
    for a in comps(la)[a0:a1]:
        for b in comps(lb)[b0:b1]:
            for c in comps(lc)[c0:c1]:
                for d in comps(ld):
                    for (cen, x) in part["combos"]:
                        q = [a, b, c, d]
                        k = CENTRES.index(cen)
                        up = list(q)
                        up[k] = inc(q[k], x)
                        T.append(tuple(up))
                        if q[k][x] > 0:
                            dn = list(q)
                            dn[k] = inc(q[k], x, -1)
                            T.append(tuple(dn))
    return T


# streaming J/K bodies of classes with at least this many integrals read each
# density value where it is used instead of preloading the six blocks (the
# (dd|dp) blocks alone are 162 doubles, i.e. 324 of the 512 registers)
JK_LAZY_D_MIN = int(os.environ.get("QX_JK_LAZY_D_MIN", "324"))

# largest SSA node count of one streaming J/K body; above it the class is
# split over ranges of a components (as the derivative parts are).  Measured
# on cluster10: splitting (dd|dp) into 3 and (dd|dd) into 6 parts made their
# J/K kernels 1.2x and 1.8x slower (recomputed recurrences cost more than
# the 5-11 KB of spills saved).  The unsplit (dd|dd) body (7614 nodes) needs
# 9.2 KB of scratch per lane, and the HSA runtime has refused to provide
# scratch above ~8 KB per lane for some dispatches (the queue aborts with
# HSA_STATUS_ERROR_OUT_OF_RESOURCES), so (dd|dd) alone is split in two.
JK_NODES = 7000


def jk_parts(l, budget=None):
    budget = budget or JK_NODES
    la, lb, lc, ld = l
    NA = ncart(la)

    def nodes(a0, a1):
        T = [(a, b, c, d) for a in comps(la)[a0:a1] for b in comps(lb) for c in comps(lc) for d in comps(ld)]
        h, v, _, _, _ = build_dag(T)
        return len(h.lines) + len(v.lines)

    for k in range(1, NA + 1):
        rs = _ranges(NA, k)
        if k == NA or all(nodes(*r) <= budget for r in rs):
            return rs


def _ranges(n, k):
    """k contiguous, nearly equal index ranges covering 0..n-1."""
    return [(i * n // k, (i + 1) * n // k) for i in range(k)]


def deriv_parts(l, budget=None):
    """Split the derivative of class l into kernels of at most `budget`
    recurrence nodes: whole class, else contiguous ranges of a components,
    else one a component with ranges of b components, else one (a, b) pair
    with ranges of c components (one (a, b, c) triple per part is the finest
    split)."""
    if budget is None:
        budget = int(_BUDGET_OVERRIDE) if _BUDGET_OVERRIDE else DERIV_BUDGET.get(tuple(l), DERIV_NODES)
    la, lb = l[0], l[1]
    NA, NB = ncart(la), ncart(lb)
    combos = [(c, x) for c in "ABC" for x in range(3)]

    NC = ncart(l[2])

    def part(ar, br, cr=(0, NC)):
        return {"combos": combos, "dinv": True, "a": ar, "b": br, "c": cr}

    def nodes(p):
        h, v, _, _, _ = build_dag(deriv_targets(l, p))
        return len(h.lines) + len(v.lines)

    whole = part((0, NA), (0, NB))
    if nodes(whole) <= budget:
        return [whole]
    for k in range(2, NA + 1):
        ps = [part(r, (0, NB)) for r in _ranges(NA, k)]
        if all(nodes(p) <= budget for p in ps):
            return ps
    for k in range(1, NB + 1):
        ps = [part((a, a + 1), r) for a in range(NA) for r in _ranges(NB, k)]
        if all(nodes(p) <= budget for p in ps):
            return ps
    # one (a, b) pair per part is still above the budget: c ranges, as many
    # as the class's csplit allows (1 keeps the (a, b) parts)
    kc = min(deriv_csplit(l), NC)
    return [part((a, a + 1), (b, b + 1), r) for a in range(NA) for b in range(NB) for r in _ranges(NC, kc)]


# Largest SSA node count of one tangent (forward-mode) derivative part.  Dual
# values double the live set: a (dd|dd) part over one (a, b) pair and all c
# (3727 nodes) needs 14.3 KB of scratch per lane, which the runtime refused
# to provide (HSA_STATUS_ERROR_OUT_OF_RESOURCES at dispatch); one c component
# per part (1860 nodes) needs 7.2 KB, no more than the plain derivative
# kernels that run.  Parts above this cap are split over single c components.
TANGENT_NODES = 2000


def tangent_parts(l):
    """Derivative parts for the tangent kernels: the class's derivative parts,
    each still above TANGENT_NODES split into one part per c component."""
    out = []
    for part in deriv_parts(l):
        h, v, _, _, _ = build_dag(deriv_targets(l, part))
        c0, c1 = part.get("c", (0, ncart(l[2])))
        if len(h.lines) + len(v.lines) <= TANGENT_NODES or c1 - c0 == 1:
            out.append(part)
        else:
            out += [dict(part, c=(c, c + 1)) for c in range(c0, c1)]
    return out


def emit_setup_tangent(L, Lt):
    """Set-up of a tangent body: every geometric quantity is a qx_dual whose
    derivative part follows the atoms along A_.e.tv (exponents are fixed)."""
    L.append("  const double p = A_.pp.p[u1], q = A_.pp.p[u2];")
    L.append("  const QxTanPair tb_ = qx_tan_pair(A_, P1, u1), tk_ = qx_tan_pair(A_, P2, u2);")
    L.append("  const qx_dual Kab = tb_.K, Kcd = tk_.K;")
    L.append("  const qx_dual Px = tb_.P[0], Py = tb_.P[1], Pz = tb_.P[2];")
    L.append("  const qx_dual Qx = tk_.P[0], Qy = tk_.P[1], Qz = tk_.P[2];")
    L.append("  const qx_dual Ax = tb_.A[0], Ay = tb_.A[1], Az = tb_.A[2];")
    L.append("  const qx_dual Cx = tk_.A[0], Cy = tk_.A[1], Cz = tk_.A[2];")
    L.append("  const R ABx = tb_.AB[0], ABy = tb_.AB[1], ABz = tb_.AB[2];")
    L.append("  const R CDx = tk_.AB[0], CDy = tk_.AB[1], CDz = tk_.AB[2];")
    L.append("  const double pq = p + q, opq = 1.0 / pq;")
    L.append("  const qx_dual Wx = (p * Px + q * Qx) * opq, Wy = (p * Py + q * Qy) * opq, "
             "Wz = (p * Pz + q * Qz) * opq;")
    L.append("  const qx_dual PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;")
    L.append("  const double rho = p * q * opq;")
    L.append("  const qx_dual T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
    L.append(f"  qx_dual F[{Lt + 1}];")
    L.append(f"  qx_boys_tan<{Lt}>(T, A_.boys, F);")
    L.append("  const qx_dual pref = (QX_TWO_PI_5_2 / (p * q * sqrt(pq))) * (Kab * Kcd);")
    for m in range(Lt + 1):
        L.append(f"  const R b{m} = pref * F[{m}];")


def emit_setup(L, Lt, deriv_w):
    L.append("  const double p = A_.pp.p[u1], Kab = A_.pp.K[u1];")
    L.append("  const double Px = A_.pp.Px[u1], Py = A_.pp.Py[u1], Pz = A_.pp.Pz[u1];")
    L.append("  const double q = A_.pp.p[u2], Kcd = A_.pp.K[u2];")
    L.append("  const double Qx = A_.pp.Px[u2], Qy = A_.pp.Py[u2], Qz = A_.pp.Pz[u2];")
    L.append("  const double Ax = A_.pr.Ax[P1], Ay = A_.pr.Ay[P1], Az = A_.pr.Az[P1];")
    L.append("  const double Cx = A_.pr.Ax[P2], Cy = A_.pr.Ay[P2], Cz = A_.pr.Az[P2];")
    L.append("  const R ABx = (R)A_.pr.ABx[P1], ABy = (R)A_.pr.ABy[P1], ABz = (R)A_.pr.ABz[P1];")
    L.append("  const R CDx = (R)A_.pr.ABx[P2], CDy = (R)A_.pr.ABy[P2], CDz = (R)A_.pr.ABz[P2];")
    L.append("  const double pq = p + q, opq = 1.0 / pq;")
    L.append("  const double Wx = (p * Px + q * Qx) * opq, Wy = (p * Py + q * Qy) * opq, Wz = (p * Pz + q * Qz) * opq;")
    L.append("  const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;")
    L.append("  const double rho = p * q * opq;")
    L.append("  const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
    L.append(f"  double F[{Lt + 1}];")
    L.append(f"  qx_boys<{Lt}>(T, A_.boys, F);")
    L.append("  const double pref = QX_TWO_PI_5_2 * Kab * Kcd / (p * q * sqrt(pq));")
    for m in range(Lt + 1):
        L.append(f"  const R b{m} = (R)(pref * F[{m}]);")


def emit_vrr_consts(L, vrr_lines):
    used = "\n".join(vrr_lines)
    for ax in AX:
        if f"PA{ax}" in used:
            L.append(f"  const R PA{ax} = (R)(P{ax} - A{ax});")
        if f"WP{ax}" in used:
            L.append(f"  const R WP{ax} = (R)(W{ax} - P{ax});")
        if f"QC{ax}" in used:
            L.append(f"  const R QC{ax} = (R)(Q{ax} - C{ax});")
        if f"WQ{ax}" in used:
            L.append(f"  const R WQ{ax} = (R)(W{ax} - Q{ax});")
    if "oo2p " in used:
        L.append("  const R oo2p = (R)(0.5 / p);")
    if "oo2q " in used:
        L.append("  const R oo2q = (R)(0.5 / q);")
    if "oo2pq " in used:
        L.append("  const R oo2pq = (R)(0.5 * opq);")
    if " rp " in used:
        L.append("  const R rp = (R)(q * opq);")
    if " rq " in used:
        L.append("  const R rq = (R)(p * opq);")


def _head(L, fn, sig, Lt, v, vn, ret="void", inline="__forceinline__", tangent=False):
    if tangent:
        L.append(f"__device__ {inline} {ret} {fn}(const QxArgs& A_, int P1, int P2, int u1, int u2{sig}) {{")
        L.append("  using R = qx_dual;")
        emit_setup_tangent(L, Lt)
    else:
        L.append(f"template <typename R>\n__device__ {inline} {ret} {fn}(const QxArgs& A_, int P1, int P2, "
                 f"int u1, int u2{sig}) {{")
        emit_setup(L, Lt, False)
    emit_vrr_consts(L, v.lines)
    L += ["  " + x for x in v.lines]
    for (a, c), name in sorted(vn.items()):
        L.append(f"  const R s_{a[0]}{a[1]}{a[2]}_{c[0]}{c[1]}{c[2]} = {name};")


def emit_eri_core(l):
    """Two bodies per class: qx_core_e* writes the integrals to an array (the
    store / tensor-scatter / Schwarz epilogues), qx_jk_e* contracts each
    integral into the J' and K' accumulators as soon as it is formed."""
    la, lb, lc, ld = l
    nn = (ncart(la), ncart(lb), ncart(lc), ncart(ld))
    NA, NB, NC, ND = nn
    N = NA * NB * NC * ND
    T = [(a, b, c, d) for a in comps(la) for b in comps(lb) for c in comps(lc) for d in comps(ld)]
    h, v, vn, outs, pos = build_dag(T)
    Lt = sum(l)
    nm = "".join(map(str, l))
    # array form
    L = []
    _head(L, f"qx_core_e{nm}", f", R (&o)[{N}]", Lt, v, vn)
    groups = [(pos[k], [f"  o[{k}] = {outs[k]};"]) for k in range(N)]
    L += ["  " + x if not x.startswith("  ") else x for x in interleave(h.lines, groups)]
    L.append("}")
    stats = dict(vrr=len(v.lines), vrr_flops=v.nflop, hrr=len(h.lines), hrr_flops=h.nflop, L=Lt, N=N,
                 flops=core_flops(Lt, v.nflop, h.nflop))
    # streaming J/K form, one body per range of a components
    jks = []
    for pi, (a0, a1) in enumerate(jk_parts(l)):
        Tp = [t for t in T if comps(la).index(t[0]) in range(a0, a1)]
        hp, vp, vnp, outp, posp = build_dag(Tp)
        J = []
        _head(J, f"qx_jk_e{nm}_{pi}", ", double (&jab)[%d], double (&jcd)[%d], double (&kac)[%d], double (&kad)[%d], "
              "double (&kbc)[%d], double (&kbd)[%d]" % (NA * NB, NC * ND, NA * NC, NA * ND, NB * NC, NB * ND),
              Lt, vp, vnp)
        lazy = N >= JK_LAZY_D_MIN
        emit_dblocks(J, nn, arrays=not lazy)

        def dv(blk, r, c):
            """density value: a preloaded block entry, or a load at the use"""
            if not lazy:
                nc_ = {"Dab": NB, "Dcd": ND, "Dac": NC, "Dad": ND, "Dbc": NC, "Dbd": ND}[blk]
                return f"{blk}[{r * nc_ + c}]"
            row, col = {"Dab": ("ia_", "ib_"), "Dcd": ("ic_", "id_"), "Dac": ("ia_", "ic_"),
                        "Dad": ("ia_", "id_"), "Dbc": ("ib_", "ic_"), "Dbd": ("ib_", "id_")}[blk]
            return f"Dm_[(size_t)({row} + {r}) * n_ + {col} + {c}]"

        groups = []
        k = 0
        for i in range(a0, a1):
            for j in range(NB):
                for kk in range(NC):
                    for ll in range(ND):
                        ij, kl, ik, il, jk, jl = (i * NB + j, kk * ND + ll, i * NC + kk, i * ND + ll, j * NC + kk,
                                                  j * ND + ll)
                        groups.append((posp[k], [
                            f"  {{ const double v_ = (double){outp[k]};",
                            f"    jab[{ij}] += v_ * {dv('Dcd', kk, ll)}; jcd[{kl}] += v_ * {dv('Dab', i, j)};",
                            f"    kac[{ik}] += v_ * {dv('Dbd', j, ll)}; kad[{il}] += v_ * {dv('Dbc', j, kk)};",
                            f"    kbc[{jk}] += v_ * {dv('Dad', i, ll)}; kbd[{jl}] += v_ * {dv('Dac', i, kk)}; }}"]))
                        k += 1
        J += ["  " + x if not x.startswith("  ") else x for x in interleave(hp.lines, groups)]
        J.append("}")
        # J/K epilogue: 6 fused multiply-adds per integral
        jks.append(("\n".join(J), (a0, a1), core_flops(Lt, vp.nflop, hp.nflop) + 12 * k))
    return "\n".join(L), jks, stats


def emit_deriv_core(l, pi, part):
    """Two bodies per derivative part: qx_core_d* writes every derivative
    block to an array (tensor scatter), qx_gr_d* contracts each derivative
    integral with the density pair product as soon as it is formed (the
    adjoint gradient), so no derivative block is ever held whole."""
    la, lb, lc, ld = l
    nn = (ncart(la), ncart(lb), ncart(lc), ncart(ld))
    NA, NB, NC_, ND = nn
    N = NA * NB * NC_ * ND
    combos, dinv = part["combos"], part["dinv"]
    a0, a1 = part["a"]
    b0, b1 = part["b"]
    c0, c1 = part.get("c", (0, ncart(lc)))
    T = deriv_targets(l, part)
    h, v, vn, outs, pos = build_dag(T)
    Lt = sum(l) + 1
    nm = "".join(map(str, l))
    NC = len(combos)
    wexpr = {"A": "(R)A_.pp.a2[u1]", "B": "(R)A_.pp.b2[u1]", "C": "(R)A_.pp.a2[u2]", "D": "(R)A_.pp.b2[u2]"}
    wdecl = [f"  const R w{c} = {wexpr[c]};" for c in sorted(set(cc for cc, _ in combos))]
    # per element: expressions of each combo and the position after which all exist
    elems = []
    ti = 0
    for ia, a in enumerate(comps(la)[a0:a1], a0):
        for ib, b in enumerate(comps(lb)[b0:b1], b0):
            for ic, c in enumerate(comps(lc)[c0:c1], c0):
                for id_, d in enumerate(comps(ld)):
                    e = ((ia * NB + ib) * NC_ + ic) * ND + id_
                    le = ((((ia - a0) * (b1 - b0) + (ib - b0)) * (c1 - c0)) + (ic - c0)) * ND + id_
                    q = [a, b, c, d]
                    exprs = []
                    last = 0
                    for ci, (cn, x) in enumerate(combos):
                        k = CENTRES.index(cn)
                        up = outs[ti]
                        last = max(last, pos[ti])
                        ti += 1
                        n = q[k][x]
                        dn = None
                        if n > 0:
                            dn = outs[ti]
                            last = max(last, pos[ti])
                            ti += 1
                        exprs.append(f"w{cn} * {up}" + (f" - R({n}) * {dn}" if dn else ""))
                    elems.append((e, le, last, exprs))
    # array form: dv holds only this part's elements (local index le)
    NEL = len(elems)
    L = []
    _head(L, f"qx_core_d{nm}_{pi}", f", R (&dv)[{NC * NEL}]", Lt, v, vn)
    L += wdecl
    groups = [(last, [f"  dv[{ci * NEL + le}] = {ex};" for ci, ex in enumerate(exprs)])
              for (e, le, last, exprs) in elems]
    L += ["  " + x if not x.startswith("  ") else x for x in interleave(h.lines, groups)]
    L.append("}")
    # tangent array form: the derivative part of each dual value, i.e. the
    # second derivative along A_.e.tv
    Tn = []
    _head(Tn, f"qx_core_t{nm}_{pi}", f", double (&dv)[{NC * NEL}]", Lt, v, vn, tangent=True)
    Tn += wdecl
    groups = [(last, [f"  dv[{ci * NEL + le}] = qx_tan({ex});" for ci, ex in enumerate(exprs)])
              for (e, le, last, exprs) in elems]
    Tn += ["  " + x if not x.startswith("  ") else x for x in interleave(h.lines, groups)]
    Tn.append("}")
    # streaming adjoint form
    G = []
    _head(G, f"qx_gr_d{nm}_{pi}", f", double (&g)[{NC}]", Lt, v, vn)
    G += wdecl
    emit_dblocks(G, nn)
    groups = []
    for (e, le, last, exprs) in elems:
        i, rem = divmod(e, NB * NC_ * ND)
        j, rem = divmod(rem, NC_ * ND)
        kk, ll = divmod(rem, ND)
        ij, kl, ik, il, jk, jl = i * NB + j, kk * ND + ll, i * NC_ + kk, i * ND + ll, j * NC_ + kk, j * ND + ll
        st = [f"  {{ const double G_ = 2.0 * Dab[{ij}] * Dcd[{kl}] - 0.5 * (Dac[{ik}] * Dbd[{jl}] + "
              f"Dad[{il}] * Dbc[{jk}]);"]
        st += [f"    g[{ci}] += G_ * (double)({ex});" for ci, ex in enumerate(exprs)]
        st.append("  }")
        groups.append((last, st))
    G += ["  " + x if not x.startswith("  ") else x for x in interleave(h.lines, groups)]
    G.append("}")
    cmask = sum(1 << CENTRES.index(c) for c in set(cc for cc, _ in combos))
    xmask = sum(1 << x for x in set(x for _, x in combos))
    nel = len(elems)
    stats = dict(part=pi, a=part["a"], b=part["b"], c=(c0, c1), dinv=dinv, vrr=len(v.lines), vrr_flops=v.nflop,
                 hrr=len(h.lines), hrr_flops=h.nflop, L=Lt, N=N, nel=nel, ncomb=NC,
                 flops=core_flops(Lt, v.nflop, h.nflop) + nel * (8 + 5 * NC))
    return "\n".join(L), "\n".join(G), "\n".join(Tn), stats, (cmask, xmask, dinv, NC, a0, a1, b0, b1, c0, c1)


def core_flops(L, vrr_flops, hrr_flops):
    """Algorithmic FP operations of one primitive-quartet body (mul and add
    count 1, fused multiply-add 2): 30 for the pair/quartet setup, 12 + 4L
    for the Boys function (Taylor step plus downward recursion), then the
    recurrences as emitted."""
    return 30 + 12 + 4 * L + vrr_flops + hrr_flops


def canonical_classes(lmax):
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    out = []
    for i, (la, lb) in enumerate(pairs):
        for j, (lc, ld) in enumerate(pairs):
            if j <= i:
                out.append((la, lb, lc, ld))
    return pairs, out


# most derivative parts one adjoint-gradient launch (and source file) covers.
# Grouping several parts under blockIdx.y would fill the GPU for classes with
# few quartets, but hipcc's time on such a file grows far faster than the
# number of parts (12 (dd|dd) parts: 24 minutes inlined, 12 minutes as
# separate functions, against ~7 s per part alone), so every part keeps its
# own kernel (the launch table and ABI still speak of groups of one).
GROUP_PARTS = 1


def part_groups(nparts):
    """Contiguous, nearly equal groups of at most GROUP_PARTS parts."""
    ng = -(-nparts // GROUP_PARTS)
    return [list(range(i * nparts // ng, (i + 1) * nparts // ng)) for i in range(ng)]


def emit_class_files(cls, outdir, fp32):
    """Writes, per class: cls_<l>_e.hip (array-form ERI kernel),
    cls_<l>_j<k>.hip (streaming J/K kernel per range of a), per derivative part
    cls_<l>_d<k>.hip (array form) and cls_<l>_t<k>.hip (tangent array form,
    second derivatives) and, per group of parts, cls_<l>_g<k>.hip
    (streaming adjoint kernel, blockIdx.y = part within the group)."""
    la, lb, lc, ld = cls
    nm = "".join(map(str, cls))
    dims = f"{ncart(la)}, {ncart(lb)}, {ncart(lc)}, {ncart(ld)}"
    N = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    files, stats = [], []

    def write(name, body):
        path = os.path.join(outdir, name)
        with open(path, "w") as fh:
            fh.write("\n".join(['#include "../qx_device.h"', ""] + body + [""]))
        files.append(path)

    src_e, jks, st = emit_eri_core(cls)
    write(f"cls_{nm}_e.hip", [src_e, "", f"QX_DEFINE_ERI_KERNEL(e{nm}, {N}, {dims})"])
    for pi, (src_j, (a0, a1), _) in enumerate(jks):
        write(f"cls_{nm}_j{pi}.hip", [src_j, "", f"QX_DEFINE_JK_KERNEL(e{nm}_{pi}, {dims}, {1 if fp32 else 0}, "
                                                  f"{a0}, {a1})"])
    st["jk_flops"] = [f for _, _, f in jks]
    stats.append((cls, "eri", st))
    parts = deriv_parts(cls)
    gsrc = []
    for pi, part in enumerate(parts):
        src_d, src_g, _, st, (cmask, xmask, dinv, NC, a0, a1, b0, b1, c0, c1) = emit_deriv_core(cls, pi, part)
        assert (cmask, xmask, dinv, NC) == (7, 7, True, 9)
        args = f"{N}, {dims}, {cmask}, {xmask}, {1 if dinv else 0}, {NC}"
        write(f"cls_{nm}_d{pi}.hip", [src_d, "", f"QX_DEFINE_DERI_KERNEL(d{nm}_{pi}, {args}, {a0}, {a1}, {b0}, {b1}, "
                                                  f"{c0}, {c1})"])
        gsrc.append(src_g)
        stats.append((cls, f"deriv{pi}", st))
    tparts = tangent_parts(cls)
    for pi, part in enumerate(tparts):
        _, _, src_t, _, (cmask, xmask, dinv, NC, a0, a1, b0, b1, c0, c1) = emit_deriv_core(cls, pi, part)
        args = f"{N}, {dims}, {cmask}, {xmask}, {1 if dinv else 0}, {NC}"
        write(f"cls_{nm}_t{pi}.hip", [src_t, "", f"QX_DEFINE_DERT_KERNEL(t{nm}_{pi}, {args}, {a0}, {a1}, {b0}, {b1}, "
                                                  f"{c0}, {c1})"])
    groups = part_groups(len(parts))
    for gi, members in enumerate(groups):
        (pi,) = members
        w = deriv_waves(cls)
        attr = f" __attribute__((amdgpu_waves_per_eu({w}, {w})))" if w > 0 else ""
        kern = ["template <typename R>",
                f"__global__ __launch_bounds__(QX_BLOCK_MAX){attr} void qx_kg_d{nm}_g{gi}(QxArgs A_) {{",
                "  __shared__ double sg[QX_GRAD_LDS];",
                "  qx_grad_body(A_, sg, [&](int P1, int P2, int u1, int u2, double (&g)[QX_GRAD_NCOMB]) {",
                f"    qx_gr_d{nm}_{pi}<R>(A_, P1, P2, u1, u2, g);", "  });", "}",
                f"QX_DEFINE_GRAD_LAUNCHER(d{nm}_g{gi}, {len(members)}, {1 if fp32 else 0})"]
        write(f"cls_{nm}_g{gi}.hip", [gsrc[pi] for pi in members] + [""] + kern)
    return files, stats, len(parts), groups, len(tparts)


def main():
    outdir = sys.argv[1]
    lmax = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    fp32 = "--no-fp32" not in sys.argv
    os.makedirs(outdir, exist_ok=True)
    _, classes = canonical_classes(lmax)
    files, allstats, nparts, flops, groups, jkf, ntparts = [], [], {}, {}, {}, {}, {}
    for cls in classes:
        f, s, np_, grp, ntp = emit_class_files(cls, outdir, fp32)
        ntparts[cls] = ntp
        files += f
        allstats += s
        nparts[cls] = np_
        groups[cls] = grp
        jkf[cls] = s[0][2]["jk_flops"]
        flops[cls] = (s[0][2]["flops"], [st["flops"] for _, _, st in s[1:]], [st["ncomb"] for _, _, st in s[1:]])
    D = ["// generated: kernel launch tables indexed by canonical class id", "#pragma once",
         f"#define QX_GEN_LMAX {lmax}", f"#define QX_NUM_CLASSES {len(classes)}",
         f"#define QX_MAX_DPARTS {max(nparts.values())}",
         f"#define QX_MAX_JKPARTS {max(len(v) for v in jkf.values())}",
         "static_assert(QX_MAX_JKPARTS <= QX_JKPARTS_CAP, \"J/K part table too small\");",
         "static_assert(QX_MAX_DPARTS <= QX_DPARTS_CAP, \"derivative part table too small\");",
         f"static_assert({max(ntparts.values())} <= QX_TPARTS_CAP, \"tangent part table too small\");", ""]
    for cls in classes:
        nm = "".join(map(str, cls))
        D.append(f'extern "C" int qx_launch_e{nm}(const QxArgs*, const QxLaunch*);')
        for pi in range(len(jkf[cls])):
            D.append(f'extern "C" int qx_launchj_e{nm}_{pi}(const QxArgs*, const QxLaunch*);')
        for pi in range(nparts[cls]):
            D.append(f'extern "C" int qx_launch_d{nm}_{pi}(const QxArgs*, const QxLaunch*);')
        for pi in range(ntparts[cls]):
            D.append(f'extern "C" int qx_launcht_t{nm}_{pi}(const QxArgs*, const QxLaunch*);')
        for gi in range(len(groups[cls])):
            D.append(f'extern "C" int qx_launchg_d{nm}_g{gi}(const QxArgs*, const QxLaunch*);')
    D.append("")
    D.append("static const QxClassEntry qx_class_table[QX_NUM_CLASSES] = {")
    for cls in classes:
        nm = "".join(map(str, cls))
        dl = ", ".join(f"qx_launch_d{nm}_{pi}" for pi in range(nparts[cls]))
        tl = ", ".join(f"qx_launcht_t{nm}_{pi}" for pi in range(ntparts[cls]))
        gl = ", ".join(f"qx_launchg_d{nm}_g{gi}" for gi in range(len(groups[cls])))
        ef, pf, pc = flops[cls]
        pfs = ", ".join(f"{x:.1f}" for x in pf)
        pcs = ", ".join(str(x) for x in pc)
        gfs = ", ".join(f"{sum(pf[p] for p in g):.1f}" for g in groups[cls])
        gps = ", ".join(str(len(g)) for g in groups[cls])
        jl = ", ".join(f"qx_launchj_e{nm}_{pi}" for pi in range(len(jkf[cls])))
        jfs = ", ".join(f"{x:.1f}" for x in jkf[cls])
        D.append(f"  {{{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, qx_launch_e{nm}, {len(jkf[cls])}, {{{jl}}}, {{{jfs}}}, "
                 f"{nparts[cls]}, {{{dl}}}, {ntparts[cls]}, {{{tl}}}, {len(groups[cls])}, {{{gl}}}, {ef:.1f}, {{{pfs}}}, {{{pcs}}}, {{{gfs}}}, "
                 f"{{{gps}}}}},")
    D.append("};")
    D.append("// per class: mixed-precision J/K and gradient passes may use FP32")
    D.append("static const unsigned char qx_fp32_default[QX_NUM_CLASSES][2] = {")
    for cls in classes:
        D.append(f"  {{{int(cls in FP32_CLASSES['jk'])}, {int(cls in FP32_CLASSES['grad'])}}},")
    D.append("};")
    D.append("// per class: the adjoint-gradient kernels reduce wave-uniform centres across the wavefront")
    D.append("static const unsigned char qx_gwred_default[QX_NUM_CLASSES] = {")
    D.append("  " + ", ".join(str(int(cls in GRAD_WAVE_REDUCE)) for cls in classes))
    D.append("};")
    D.append("// per class: copies of the gradient accumulators its adjoint blocks flush into")
    D.append("static const int qx_gcopy_default[QX_NUM_CLASSES] = {")
    D.append("  " + ", ".join(str(GRAD_COPIES.get(cls, 64)) for cls in classes))
    D.append("};")
    D.append("// per class: threads per block of the streaming J/K and gradient kernels")
    D.append("static const int qx_tile_default[QX_NUM_CLASSES][2] = {")
    for cls in classes:
        D.append(f"  {{{TILES['jk'].get(cls, 128)}, {TILES['grad'].get(cls, 128)}}},")
    D.append("};")
    with open(os.path.join(outdir, "qx_classes.inc"), "w") as fh:
        fh.write("\n".join(D) + "\n")
    with open(os.path.join(outdir, "stats.txt"), "w") as fh:
        for cls, kind, st in allstats:
            fh.write(f"{cls} {kind} {st}\n")
    print("\n".join(files))


if __name__ == "__main__":
    main()
