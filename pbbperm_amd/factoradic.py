"""Factoradic work-unit codec (distributed/src/fact_work.cpp:47-86,
distributed/src/weights.cpp:10-26): an IVM interval is a pair of digit
vectors pos/end, digit i in [0, n-1-i], whose value is sum_i d_i * (n-1-i)!.
The reference uses GMP bigints; Python integers are exact at any n.
"""
from __future__ import annotations

from math import factorial
from typing import List, Sequence, Tuple

import numpy as np


def weights(n: int) -> List[int]:
    """Place values (n-1)!, (n-2)!, ..., 0!, as weights::depth[1..n]."""
    return [factorial(n - 1 - i) for i in range(n)]


def to_int(digits: Sequence[int]) -> int:
    """fact_work::VectToBigint."""
    n = len(digits)
    w = weights(n)
    v = 0
    for i, d in enumerate(digits):
        d = int(d)
        if d < 0 or d > n - 1 - i:
            raise ValueError(f"digit {i} = {d} outside [0, {n - 1 - i}]")
        v += d * w[i]
    return v


def to_digits(value: int, n: int) -> np.ndarray:
    """fact_work::BigintToVect (quotients by (n-1)!, (n-2)!, ...)."""
    if value < 0 or value >= factorial(n):
        raise ValueError("value outside [0, n!-1]")
    out = np.zeros(n, dtype=np.intc)
    r = int(value)
    for i, w in enumerate(weights(n)):
        out[i], r = divmod(r, w)
    return out


def full_interval(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """[0, n!-1]: pos = 0...0, end = (n-1, ..., 0) (sequentialbb.cpp:72-75)."""
    return np.zeros(n, np.intc), np.arange(n - 1, -1, -1, dtype=np.intc)


def split(n: int, parts: int, lo: int = 0, hi: int | None = None):
    """Cut [lo, hi] into `parts` contiguous closed intervals of near-equal
    length (works::init with initialWork=3 cuts [0, n!] into nProc parts).
    Returns (pos[k, n], end[k, n])."""
    hi = factorial(n) - 1 if hi is None else hi
    total = hi - lo + 1
    parts = max(1, min(parts, total))
    pos, end = [], []
    start = lo
    for k in range(parts):
        size = total // parts + (1 if k < total % parts else 0)
        pos.append(to_digits(start, n))
        end.append(to_digits(start + size - 1, n))
        start += size
    return np.stack(pos), np.stack(end)


def length(pos: Sequence[int], end: Sequence[int]) -> int:
    """Number of leaves in the closed interval [pos, end] (0 if empty)."""
    a, b = to_int(pos), to_int(end)
    return max(0, b - a + 1)


def to_decimal_text(pos: Sequence[int], end: Sequence[int], ident: int) -> str:
    """One interval in the master's decimal text form (`id begin end`)."""
    return f"{ident} {to_int(pos)} {to_int(end)}"
