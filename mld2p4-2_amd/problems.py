"""Synthetic test problems (host side), vectorised with numpy.

pde3d_poisson: the 7-point finite-difference Poisson matrix of the
reference's 3-D driver, tests/pdegen/mld_d_pde3d.f90:98-412 with
a1=a2=a3=1, b1=b2=b3=c=0:
  h = 1/(idim+2) (:165); global row (ix-1)*idim^2 + (iy-1)*idim + iz, z fastest
  (:258-268); per row the entries x-1, y-1, z-1, centre, z+1, y+1, x+1 with
  values -1/h^2 and 6/h^2 (:278-336, :417-458); boundary neighbours dropped
  (Dirichlet data goes to the right-hand side, which the benchmark replaces
  by all ones).
Returns CSR arrays (irp int64, ja int32, val float64), columns ascending.
"""
from __future__ import annotations

import numpy as np


def pde3d_poisson(idim: int):
    n = idim ** 3
    h = 1.0 / (idim + 2)
    sq = h * h
    off = -1.0 / sq - 0.0 / (2.0 * h)
    cen = 2.0 * (1.0 + 1.0 + 1.0) / sq + 0.0
    g = np.arange(n, dtype=np.int64)
    x = g // (idim * idim)
    y = (g // idim) % idim
    z = g % idim
    d2 = idim * idim
    # neighbour offsets in ascending column order and their validity masks
    offs = [(-d2, x > 0), (-idim, y > 0), (-1, z > 0), (0, None), (1, z < idim - 1), (idim, y < idim - 1),
            (d2, x < idim - 1)]
    counts = np.ones(n, dtype=np.int64)
    for o, m in offs:
        if m is not None:
            counts += m
    irp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=irp[1:])
    nnz = int(irp[-1])
    ja = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    pos = irp[:-1].copy()
    for o, m in offs:
        if m is None:
            ja[pos] = g
            val[pos] = cen
            pos += 1
        else:
            idx = pos[m]
            ja[idx] = (g + o)[m]
            val[idx] = off
            pos += m
    return irp, ja, val


def csr_matvec(irp, ja, val, x):
    """Host y = A x (numpy), used for residual checks."""
    import scipy.sparse as sp

    n = len(irp) - 1
    return sp.csr_matrix((val, ja, irp), shape=(n, len(x))) @ x
