"""Exact single-precision arithmetic helpers for the oracle pins.

fmaf32 is a vectorised, correctly rounded float32 fused multiply-add written
from its definition (round-to-nearest-even of a*b + c computed exactly),
independent of the C library's fmaf that the oracle calls:
- a*b of two floats is exact in float64 (48 significant bits);
- TwoSum gives s = fl64(p + c) and its exact error e, so p + c = s + e;
- every float32 rounding boundary (a midpoint of two neighbouring floats) is
  a float64 value, and |e| <= ulp64(s)/2, so s + e rounds to the same float
  as s unless s is itself a midpoint; there the sign of e decides, and e = 0
  is a true tie, rounded to even (what float64 -> float32 conversion does).
"""
import numpy as np


def fmaf32(a, b, c):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    c = np.asarray(c, np.float32)
    p = a.astype(np.float64) * b.astype(np.float64)
    c64 = c.astype(np.float64)
    s = p + c64
    bv = s - p
    e = (p - (s - bv)) + (c64 - bv)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    toward = np.where(s > r64, np.inf, -np.inf).astype(np.float32)
    nb = np.nextafter(r, toward)
    mid = (r64 + nb.astype(np.float64)) / 2.0
    at_mid = (s != r64) & (s == mid)
    # at a midpoint, s rounded to r (even); the exact value moves past it when
    # e points from r toward nb
    up = at_mid & (np.sign(e) == np.sign(nb.astype(np.float64) - r64)) & (e != 0)
    return np.where(up, nb, r).astype(np.float32)


def dot_fma_chain(x, planes):
    """dot[i][s] = sequential fmaf over j of x[i][j] * planes[s][j] from 0.0f,
    in index order (the FP32 reference order of simHash.cu:22-47 and
    simHashJL.cuh:34-51)."""
    x = np.asarray(x, np.float32)
    planes = np.asarray(planes, np.float32)
    acc = np.zeros((x.shape[0], planes.shape[0]), np.float32)
    for j in range(x.shape[1]):
        acc = fmaf32(x[:, j][:, None], planes[:, j][None, :], acc)
    return acc
