"""The parity comparator shared by the GPU tests and the smoke check.

Two FP64 arrays agree when, position by position:
  * both are the same value (this covers infinities of the same sign), or
  * both are NaN, or
  * both are finite and |a - b| <= rtol * max(|a|, |b|).
An infinity matches only an infinity of the same sign, and a NaN only a
NaN.  Path counts can wrap to 0 mod 2^32, after which (delta + 1) / 0 is
infinite and inf * 0 is NaN; those positions must agree exactly.
"""
from __future__ import annotations

import numpy as np

RTOL = 1e-9


def close(a, b, rtol: float = RTOL):
    """(all agree, worst relative error over the finite pairs)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        return False, float("inf")
    fin = np.isfinite(a) & np.isfinite(b)
    with np.errstate(invalid="ignore", over="ignore"):
        scale = np.maximum(np.abs(a), np.abs(b))
        diff = np.where(fin, np.abs(a - b), 0.0)
    ok = (a == b) | (np.isnan(a) & np.isnan(b)) | (fin & (diff <= rtol * scale))
    err = float((diff / np.where(scale > 0, scale, 1.0))[fin].max()) if fin.any() else 0.0
    return bool(ok.all()), err
