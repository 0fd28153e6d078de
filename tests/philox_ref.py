"""Host restatement, for tests only, of two pieces of the library's own
Philox-mode algorithm (not of the reference, whose draws are numpy's):

* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
  easy as 1, 2, 3", SC'11), checked against the published known-answer
  vectors in tests/test_host.py;
* the frame-end detector noise of the Philox path: per pixel, dark counts
  Poisson(lambda) by sequential inversion, then afterpulses
  Binomial(counts after the dark counts, p_ap) by inversion (BINV), each
  count into a uniform bin, all keyed by (frame, pixel, count index, tag).

The GPU test compares the device's frame noise with this draw for draw.
"""

from __future__ import annotations

import math
from typing import Tuple

import numpy as np

MASK = 0xFFFFFFFF
_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
TAG_DARK, TAG_AFTERPULSE = 15, 16


def philox4x32(ctr, k0: int, k1: int) -> Tuple[int, int, int, int]:
    x, y, z, w = (int(v) & MASK for v in ctr)
    k0 &= MASK
    k1 &= MASK
    for _ in range(10):
        p0 = _M0 * x
        p1 = _M1 * z
        x, y, z, w = ((p1 >> 32) ^ y ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ w ^ k1) & MASK, p0 & MASK
        k0 = (k0 + _W0) & MASK
        k1 = (k1 + _W1) & MASK
    return x, y, z, w


def u01d(a: int, b: int) -> float:
    """53-bit double in [0, 1) from two words."""
    return float(((a >> 5) << 26) | (b >> 6)) * (1.0 / 9007199254740992.0)


def poisson_inv(lam: float, u: float) -> int:
    p = math.exp(-lam)
    F = p
    k = 0
    while u > F and k < 1000:
        k += 1
        p *= lam / k
        F += p
    return k


def binomial_inv(n: int, p: float, u: float) -> int:
    """BINV (Kachitvichyanukul & Schmeiser 1988) for n p <= 30."""
    if n <= 0 or p <= 0.0:
        return 0
    assert n * p <= 30.0, "the normal branch is not restated"
    q = 1.0 - p
    s = p / q
    a = (n + 1) * s
    r = q ** n
    k = 0
    while u > r and k < n:
        u -= r
        k += 1
        r *= (a / k - s)
    return k


def frame_noise(flim: np.ndarray, lam: float, p_ap: float, seed: int, frame: int):
    """flim: [n_pixels, n_bins] uint32 before the frame end.  Returns the
    histogram after dark counts and afterpulses and the two counts."""
    k0 = (seed & MASK) ^ 0x2545F491
    k1 = (seed >> 32) & MASK
    out = flim.astype(np.int64).copy()
    npix, nb = out.shape
    nd = na = 0
    for q in range(npix):
        lo, hi = q & MASK, (q >> 32) & MASK
        tot = int(flim[q].sum())
        if lam >= 1e-10:
            r = philox4x32((frame, lo, hi, TAG_DARK), k0, k1)
            k = poisson_inv(lam, u01d(r[0], r[1]))
            for i in range(k):
                rb = philox4x32((frame, lo, (hi + 0x10000 * (i + 1)) & MASK, TAG_DARK), k0, k1)
                out[q, int(u01d(rb[0], rb[1]) * nb) % nb] += 1
            nd += k
            tot += k
        if p_ap > 0.0 and tot:
            r = philox4x32((frame, lo, hi, TAG_AFTERPULSE), k0, k1)
            k = binomial_inv(tot, p_ap, u01d(r[0], r[2]))
            for i in range(k):
                rb = philox4x32((frame, lo, (hi + 0x10000 * (i + 1)) & MASK, TAG_AFTERPULSE), k0,
                                k1)
                out[q, int(u01d(rb[0], rb[1]) * nb) % nb] += 1
            na += k
    return out.astype(np.uint32), nd, na
