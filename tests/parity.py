"""Comparison helpers for GPU-vs-oracle parity (tolerances stated here).

FP64 fields must agree to |g - c| <= RTOL * max(|g|, |c|) + ATOL[field];
integer fields bit for bit.  RTOL = 1e-10 is the parity bar; the absolute
floors sit at the reference's own zeroing threshold (minStorVol = 1e-15 km3,
routing.h:24) for km3 storages and at 1e-12 mm for the vertical storages.
"""
import numpy as np

RTOL = 1e-10
KM3 = ["groundwater", "loc_lake", "loc_wetl", "glo_lake", "glo_wetl", "reservoir", "river"]
ATOL = {k: 1e-14 for k in KM3}
ATOL.update({"discharge": 1e-14})


def atol(name):
    return ATOL.get(name, 1e-12)


def compare(gpu, cpu, fields=None, rtol=RTOL, label=""):
    """Returns a list of (field, worst_excess, index) for fields outside tolerance."""
    bad = []
    for k in (fields or [k for k in cpu if k != "started"]):
        g = np.asarray(gpu[k])
        c = np.asarray(cpu[k])
        if g.dtype.kind in "iu":
            if not np.array_equal(g, c):
                idx = np.flatnonzero(g.ravel() != c.ravel())
                bad.append((label + k, "int mismatch", int(idx[0]), len(idx)))
            continue
        g = g.astype(np.float64).ravel()
        c = c.astype(np.float64).ravel()
        both_nan = np.isnan(g) & np.isnan(c)
        diff = np.abs(g - c)
        tol = rtol * np.maximum(np.abs(g), np.abs(c)) + atol(k)
        over = (diff > tol) & ~both_nan
        if over.any():
            i = int(np.argmax(np.where(over, diff / np.maximum(tol, 1e-300), 0)))
            bad.append((label + k, float(diff[i]), i, int(over.sum()), float(g[i]), float(c[i])))
    return bad


def assert_close(gpu, cpu, fields=None, rtol=RTOL, label=""):
    bad = compare(gpu, cpu, fields, rtol, label)
    assert not bad, bad[:8]
