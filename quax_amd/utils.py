"""Derivative index contract of the ERI path (quax/utils.py:39-70)."""
from __future__ import annotations

import itertools
import math

import numpy as np


def how_many_derivs(k: int, n: int) -> int:
    """Unique Cartesian derivatives of order n for k atoms: C(3k+n-1, n)
    (utils.py:39-47)."""
    return math.comb(3 * k + n - 1, n) if n >= 0 else 0


def get_deriv_vec_idx(deriv_vec) -> int:
    """Index of a derivative vector in the flattened upper hypertriangle of
    the Cartesian derivative tensor (utils.py:49-70): the rank of its sorted
    multi-index among combinations with replacement in lexicographic order.
    Computed by counting instead of enumerating."""
    dv = np.asarray(deriv_vec, dtype=int).reshape(-1)
    if (dv < 0).any():
        raise ValueError("deriv_vec entries must be non-negative")
    dim = dv.shape[0]
    idx = [i for i in range(dim) for _ in range(dv[i])]
    order = len(idx)
    rank = 0
    prev = 0
    for pos, v in enumerate(idx):
        remaining = order - pos - 1
        for smaller in range(prev, v):
            # combinations of `remaining` elements drawn from [smaller, dim)
            rank += math.comb(dim - smaller + remaining - 1, remaining)
        prev = v
    return rank


def generate_multi_index_lookup(nparams: int, deriv_order: int):
    return [list(c) for c in itertools.combinations_with_replacement(range(nparams), deriv_order)]
