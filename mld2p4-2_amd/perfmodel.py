"""Algorithmic (compulsory) HBM bytes of the cycle steps.

Every array an operation must touch is counted once: matrix values and
column indices (s_v + 4 bytes per stored entry of the unpadded operator),
row pointers for operators stored as CSR, and each vector read or written
once.  ELL padding and cache re-reads are not algorithmic and are not
counted, so achieved GB/s = bytes / kernel time is a lower bound on the
bandwidth the kernel actually moved.
"""
from __future__ import annotations

DIAG_SCALE, ILU_N, L1_DIAG, DENSE_LU, F_NONE = 5, 7, 101, 102, 4


def mat_bytes(nr, nnz, sv, fmt):
    """Compulsory bytes of one pass over a stored operator: values and column
    indices (CSR also row pointers; DIA stores no indices; SELD stores 16-bit
    column steps and one 32-bit first column per row)."""
    if fmt == "DIA":
        return nnz * sv
    if fmt == "SELD":
        return nnz * (sv + 2) + nr * 4
    b = nnz * (sv + 4)
    if fmt == "CSR":
        b += (nr + 1) * 4
    return b


def spmv_bytes(nr, nc, nnz, sv, sx, beta_read=False, csr=True, nb=1, fmt=None):
    """y = A x (beta_read: y = y + A x); nb columns share one pass over A."""
    fmt = fmt or ("CSR" if csr else "ELL")
    return mat_bytes(nr, nnz, sv, fmt) + nb * (nc * sx + nr * sx * (2 if beta_read else 1))


def residual_bytes(n, nnz, sv, sx, csr=False, nb=1, fmt=None):
    # r = b - A y : A, y (gathered, once), b, r
    fmt = fmt or ("CSR" if csr else "ELL")
    return mat_bytes(n, nnz, sv, fmt) + 3 * n * sx * nb


def smoother_bytes(n, solver, sweeps, nnz_nd, nnz_lu, sv, sx, beta, nd_csr=False, nb=1, nd_fmt=None):
    """One sm%apply(1, x, beta, y, sweeps, init='Z') on nb columns."""
    yread = n * sx * nb if beta else 0
    if solver in (DIAG_SCALE, L1_DIAG):
        if sweeps <= 1 or nnz_nd == 0:
            return 2 * n * sx * nb + n * sv + yread  # y = d*x
        m = mat_bytes(n, nnz_nd, sv, nd_fmt or ("CSR" if nd_csr else "ELL"))
        first = m + n * (2 * sx * nb + sv)  # D^-1(x - ND D^-1 x)
        more = (sweeps - 2) * (m + n * (3 * sx * nb + sv))
        return first + more + yread
    if solver == ILU_N:
        return nnz_lu * (sv + 4) + 2 * (n + 1) * 4 + n * sv + 4 * n * sx + yread
    if solver == DENSE_LU:
        return n * n * 8 + n * 4 + 2 * n * sx + yread
    return 2 * n * sx + yread


def vcycle_bytes(levels, sv, sx):
    """levels: list of dicts with n, nnz_a, nnz_p, nnz_r, nnz_nd, nnz_lu,
    solver, sweeps_pre, sweeps_post, csr_a (bool).  P/R at level l map
    level l-1 <-> l.  Returns (total, per-step list)."""
    steps = []
    nlev = len(levels)
    for l, L in enumerate(levels):
        n = L["n"]
        if l < nlev - 1:
            C = levels[l + 1]
            nc = C["n"]
            steps.append((l, "pre", smoother_bytes(n, L["solver"], L["sweeps_pre"], L["nnz_nd"], L.get("nnz_lu", 0),
                                                   sv, sx, 0, L.get("csr_nd", False), nd_fmt=L.get("fmt_nd"))))
            steps.append((l, "residual", residual_bytes(n, L["nnz_a"], sv, sx, L.get("csr_a", False),
                                                        fmt=L.get("fmt_a"))))
            steps.append((l + 1, "restrict", spmv_bytes(nc, n, C["nnz_r"], sv, sx, fmt=C.get("fmt_r"))))
            steps.append((l + 1, "prolong", spmv_bytes(n, nc, C["nnz_p"], sv, sx, beta_read=True, fmt=C.get("fmt_p"))))
            steps.append((l, "residual", residual_bytes(n, L["nnz_a"], sv, sx, L.get("csr_a", False),
                                                        fmt=L.get("fmt_a"))))
            steps.append((l, "post", smoother_bytes(n, L["solver"], L["sweeps_post"], L["nnz_nd"], L.get("nnz_lu", 0),
                                                    sv, sx, 1, L.get("csr_nd", False), nd_fmt=L.get("fmt_nd"))))
        else:
            steps.append((l, "coarse", smoother_bytes(n, L["solver"], L["sweeps_pre"], L["nnz_nd"],
                                                      L.get("nnz_lu", 0), sv, sx, 0)))
    return sum(s[2] for s in steps), steps


def block_vcycle_bytes(levels, sv, sx, nb, coarse_n):
    """Block V-cycle on nb columns (every operator CSR, read once per
    product for all columns); the coarsest solve is a dense coarse_n^2
    operator."""
    tot = 0
    nlev = len(levels)
    for l, L in enumerate(levels):
        n = L["n"]
        if l < nlev - 1:
            C = levels[l + 1]
            nc = C["n"]
            tot += smoother_bytes(n, L["solver"], L["sweeps_pre"], L["nnz_nd"], 0, sv, sx, 0, True, nb)
            tot += 2 * residual_bytes(n, L["nnz_a"], sv, sx, True, nb)
            tot += spmv_bytes(nc, n, C["nnz_r"], sv, sx, nb=nb)
            tot += spmv_bytes(n, nc, C["nnz_p"], sv, sx, beta_read=True, nb=nb)
            tot += smoother_bytes(n, L["solver"], L["sweeps_post"], L["nnz_nd"], 0, sv, sx, 1, True, nb)
        else:
            tot += spmv_bytes(coarse_n, coarse_n, coarse_n * coarse_n, sv, sx, nb=nb)
    return tot


def fv7_grad_bytes(ncell):
    """theta, x, lambda read and the gradient written, 8 bytes each."""
    return 32 * ncell
