"""Entry-wise value tolerance of the parity tests.

Every entry d = |a - ref| must satisfy d <= tol * max(|ref|, floor * s), with
s the largest |ref| of the entry's matrix row (CSR values, given rowptr) or of
the whole vector.  The floor keeps entries that cancel to far below their
row's scale (exact zeros of the exact problem, rounding residues, sums of
many cancelling quadrature terms) to an absolute bound of tol * floor * s
instead of a relative one: two evaluation orders of such an entry differ by
a few ulps of its terms, which are of the size of the row's large entries
(measured: up to 7e-14 s on coupling rows of the Taylor-Hood pairs).
"""
import numpy as np


def entrywise(a, ref, rowptr=None, tol=1e-12, floor=0.1):
    """(ok, worst d / max(|ref|, floor * s)) for equal-length arrays."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert a.shape == ref.shape, (a.shape, ref.shape)
    if ref.size == 0:
        return True, 0.0
    absr = np.abs(ref)
    if rowptr is not None:
        rowptr = np.asarray(rowptr, dtype=np.int64)
        counts = np.diff(rowptr)
        rowmax = np.zeros(len(counts))
        nz = counts > 0
        rowmax[nz] = np.maximum.reduceat(absr, rowptr[:-1][nz])
        s = np.repeat(rowmax, counts)
    else:
        s = absr.max()
    bound = np.maximum(absr, floor * s)
    d = np.abs(a - ref)
    ok = bool(np.all(d <= tol * bound))
    ratio = d / np.where(bound > 0, bound, 1.0)
    return ok, float(ratio.max())
