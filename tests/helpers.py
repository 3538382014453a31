"""Error metrics shared by the parity tests."""
import numpy as np


def rel(a, b):
    """Normwise relative difference |a - b| / |b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0:
        return d
    return d / nb


def max_entry_rel(a, b, floor=1e-12):
    """max |a - b| / max(|b|, floor * max|b|): entrywise relative with an absolute floor."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    scale = np.maximum(np.abs(b), floor * np.max(np.abs(b)) if b.size else 1.0)
    scale[scale == 0] = 1.0
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def rel_finite(a, b):
    """rel() over the entries where b is finite; non-finite entries must match exactly (the
    reference produces inf/nan at singular configurations, e.g. ARAP at sigma_i + sigma_j = 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    fin = np.isfinite(b)
    if not np.array_equal(np.isfinite(a), fin):
        return np.inf
    if not np.array_equal(a[~fin], b[~fin], equal_nan=True):
        return np.inf
    return rel(a[fin], b[fin])


def bsr_entry_rel(K, Kref, row_ptr, floor=1e-3):
    """Entrywise relative error of BSR blocks (nnzb, 3, 3) with a per-row floor: each scalar
    entry is compared against max(|ref|, floor * largest |ref| in its scalar row).  Entries that
    are sums of cancelling element contributions carry round-off of order eps * (row scale), so a
    bare entrywise ratio would measure cancellation, not agreement."""
    K = np.asarray(K, dtype=np.float64)
    Kref = np.asarray(Kref, dtype=np.float64)
    nv = len(row_ptr) - 1
    worst = 0.0
    for i in range(nv):
        blk = Kref[row_ptr[i]:row_ptr[i + 1]]            # (n, 3, 3)
        rowmax = np.abs(blk).max(axis=(0, 2))             # per scalar row a
        scale = np.maximum(np.abs(blk), floor * rowmax[None, :, None])
        scale[scale == 0] = 1.0
        d = np.abs(K[row_ptr[i]:row_ptr[i + 1]] - blk) / scale
        worst = max(worst, float(d.max()))
    return worst
