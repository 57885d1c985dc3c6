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


def pde2d_poisson(idim: int, single_h: bool = True):
    """5-point matrix of mld_d_gen_pde2d (tests/pdegen/mld_d_pde2d.f90:96-370)
    with a1 = a2 = 1, b1 = b2 = c = 0: rows x-major (glob_row = (ix-1)*idim +
    iy), columns (x-1, y-1, centre, y+1, x+1).  The reference forms the mesh
    size as 1.e0/(idim+2), a default-real (single precision) quotient widened
    to double (:163); single_h=False uses 1/(idim+2) in double, the mesh size
    of the reference's golden matrix tests/fileread/runs/mld_mat.mtx."""
    n = idim * idim
    h = float(np.float32(1.0) / np.float32(idim + 2)) if single_h else 1.0 / (idim + 2)
    sq = h * h
    off = -1.0 / sq - 0.0 / (2.0 * h)
    cen = 2.0 * (1.0 + 1.0) / sq + 0.0
    g = np.arange(n, dtype=np.int64)
    x = g // idim
    y = g % idim
    offs = [(-idim, x > 0), (-1, y > 0), (0, None), (1, y < idim - 1), (idim, x < idim - 1)]
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


def box3(g):
    """One pass of the 3x3x3 box filter with edge replication, applied as
    three 3-point means along the axes (separable form of the 27-point mean)."""
    for ax in range(3):
        p = np.pad(g, [(1, 1) if a == ax else (0, 0) for a in range(3)], mode="edge")
        sl = [slice(None)] * 3
        lo, mid, hi = list(sl), list(sl), list(sl)
        lo[ax], mid[ax], hi[ax] = slice(0, -2), slice(1, -1), slice(2, None)
        g = (p[tuple(lo)] + p[tuple(mid)] + p[tuple(hi)]) / 3.0
    return g


def diffusion3d(idim: int, seed: int = 2024, passes: int = 3, contrast: float = 2.0):
    """Heterogeneous diffusion, 7-point finite-volume stencil on the unit cube
    with idim^3 cells (h = 1/idim), zero Dirichlet data on the boundary.

    g ~ N(0,1) i.i.d. per cell from numpy's Philox generator seeded with
    `seed`, smoothed by `passes` box filters; k = exp(contrast*g).  Face
    coefficient between neighbours i, j: 2 k_i k_j / (k_i + k_j) (harmonic
    mean); a boundary face contributes k_i.  Row i: -c_f/h^2 off the diagonal,
    sum_f c_f / h^2 on it; columns ascending (x-1, y-1, z-1, i, z+1, y+1, x+1
    with z fastest, as the Poisson generator).  b ~ U[-1, 1] from the same
    generator.  Returns (irp, ja, val, b)."""
    rng = np.random.Generator(np.random.Philox(seed))
    g = rng.standard_normal((idim, idim, idim))
    for _ in range(passes):
        g = box3(g)
    kc = np.exp(contrast * g)
    b = rng.uniform(-1.0, 1.0, idim ** 3)
    irp, ja, val = fv7_matrix(kc)
    return irp, ja, val, b


def fv7_matrix(kc):
    """7-point finite-volume matrix of -div(k grad u) on the unit cube with
    cell coefficients kc (idim^3 array, C order, z fastest), h = 1/idim,
    harmonic face coefficients, zero Dirichlet faces contributing k_i."""
    kc = np.asarray(kc, dtype=np.float64)
    idim = int(round(kc.size ** (1.0 / 3.0)))
    kc = kc.reshape(idim, idim, idim)
    h2 = (1.0 / idim) ** 2

    def harm(a, c):
        return 2.0 * a * c / (a + c)

    n = idim ** 3
    # face coefficients towards +x, +y, +z (0 where the face is on the boundary)
    fx = np.zeros_like(kc)
    fy = np.zeros_like(kc)
    fz = np.zeros_like(kc)
    fx[:-1] = harm(kc[:-1], kc[1:])
    fy[:, :-1] = harm(kc[:, :-1], kc[:, 1:])
    fz[:, :, :-1] = harm(kc[:, :, :-1], kc[:, :, 1:])
    # diagonal: interior faces + boundary faces (k_i)
    diag = np.zeros_like(kc)
    diag += fx + fy + fz
    diag[1:] += fx[:-1]
    diag[:, 1:] += fy[:, :-1]
    diag[:, :, 1:] += fz[:, :, :-1]
    diag[0] += kc[0]
    diag[-1] += kc[-1]
    diag[:, 0] += kc[:, 0]
    diag[:, -1] += kc[:, -1]
    diag[:, :, 0] += kc[:, :, 0]
    diag[:, :, -1] += kc[:, :, -1]
    fx, fy, fz, diag = (v.reshape(n) for v in (fx, fy, fz, diag))
    gi = np.arange(n, dtype=np.int64)
    x = gi // (idim * idim)
    y = (gi // idim) % idim
    z = gi % idim
    d2 = idim * idim
    # (offset, mask, value) in ascending column order
    terms = [(-d2, x > 0, None, "xm"), (-idim, y > 0, None, "ym"), (-1, z > 0, None, "zm"), (0, None, None, "c"),
             (1, z < idim - 1, None, "zp"), (idim, y < idim - 1, None, "yp"), (d2, x < idim - 1, None, "xp")]
    vals = {
        "xm": lambda m: -np.roll(fx, d2)[m] / h2, "ym": lambda m: -np.roll(fy, idim)[m] / h2,
        "zm": lambda m: -np.roll(fz, 1)[m] / h2, "zp": lambda m: -fz[m] / h2, "yp": lambda m: -fy[m] / h2,
        "xp": lambda m: -fx[m] / h2,
    }
    counts = np.ones(n, dtype=np.int64)
    for o, m, _, _ in terms:
        if m is not None:
            counts += m
    irp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=irp[1:])
    ja = np.empty(int(irp[-1]), dtype=np.int32)
    val = np.empty(int(irp[-1]), dtype=np.float64)
    pos = irp[:-1].copy()
    for o, m, _, key in terms:
        if m is None:
            ja[pos] = gi
            val[pos] = diag / h2
            pos += 1
        else:
            idx = pos[m]
            ja[idx] = (gi + o)[m]
            val[idx] = vals[key](m)
            pos += m
    return irp, ja, val


def csr_matvec(irp, ja, val, x):
    """y = A x for CSR arrays (host, vectorised; used for true residuals)."""
    rows = np.repeat(np.arange(len(irp) - 1), np.diff(irp))
    return np.bincount(rows, weights=val * x[ja], minlength=len(irp) - 1)


def theta_problem(idim: int, seed: int = 5, sigma: float = 0.5):
    """Inverse-problem workload: per-cell log-coefficient theta ~ N(0, sigma^2)
    (Philox, `seed`), b ~ U[-1, 1] from the same generator, and a second
    field theta* from seed+1 whose solution is the target.  k = exp(theta)."""
    rng = np.random.Generator(np.random.Philox(seed))
    theta = sigma * rng.standard_normal(idim ** 3)
    b = rng.uniform(-1.0, 1.0, idim ** 3)
    rng2 = np.random.Generator(np.random.Philox(seed + 1))
    theta_star = sigma * rng2.standard_normal(idim ** 3)
    return theta, b, theta_star
