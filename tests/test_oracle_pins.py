"""Pins of the oracle (CPU restatement) against what fixes it independently:
the reference's golden vectors, closed forms, a second pure-Python
restatement, invariants and dense brute force on tiny inputs.  None of these
checks retypes the oracle's own loop: each uses an independent formulation
that a plausible slip (dropped term, wrong sign, transposed operand, wrong
sweep count) would violate.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def poisson_kron(idim):
    """3-D 7-point Poisson from Kronecker sums of integer 1-D stencils,
    divided by h^2 once (h = 1/(idim+2)): an independent construction."""
    T = sp.diags([-np.ones(idim - 1), 2 * np.ones(idim), -np.ones(idim - 1)], [-1, 0, 1], dtype=np.int64)
    I = sp.identity(idim, dtype=np.int64)
    K = sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
    h = 1.0 / (idim + 2)
    K = sp.csr_matrix(K)
    K.sum_duplicates()
    K.eliminate_zeros()
    K.sort_indices()
    return sp.csr_matrix((K.data.astype(np.float64) / (h * h), K.indices, K.indptr), shape=K.shape)


@pytest.mark.parametrize("idim", [2, 3, 5])
def test_gen_pde3d_closed_form(orc, idim):
    A = orc.gen_pde3d(idim).to_scipy()
    K = poisson_kron(idim)
    assert (A.indptr == K.indptr).all() and (A.indices == K.indices).all()
    assert (A.data == K.data).all()


def test_spmv_against_scipy_and_loop(orc):
    a = orc.gen_pde3d(6)
    A = a.to_scipy()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(A.shape[0])
    y = a.spmv(x)
    assert np.allclose(y, A @ x, rtol=1e-14, atol=1e-10)
    # stored-order sequential sum, row by row
    ref = np.array([sum(A.data[k] * x[A.indices[k]] for k in range(A.indptr[i], A.indptr[i + 1]))
                    for i in range(A.shape[0])])
    assert (y == ref).all()
    y0 = rng.standard_normal(A.shape[0])
    assert np.allclose(a.spmv(x, alpha=-1.0, beta=1.0, y=y0.copy()), y0 - A @ x, rtol=1e-14, atol=1e-9)
    assert np.allclose(a.spmv(x, trans=True), A.T @ x, rtol=1e-14, atol=1e-10)


# ---------------------------------------------------------------- aggregation
def py_dec_map_bld(A, theta):
    """Second, independent restatement of the greedy decoupled aggregation
    (mld_d_dec_map_bld.f90:148-278), plain Python on small matrices."""
    n = A.shape[0]
    d = A.diagonal()
    rows = [(A.indices[A.indptr[i]:A.indptr[i + 1]], A.data[A.indptr[i]:A.indptr[i + 1]]) for i in range(n)]

    def strong(i):
        cols, vals = rows[i]
        return [(j, v) for j, v in zip(cols, vals) if abs(v) > theta * np.sqrt(abs(d[i] * d[j]))]

    agg = [None] * n
    na = 0
    for i in range(n):
        if agg[i] is not None:
            continue
        nb = [j for j, _ in strong(i)]
        if all(agg[j] is None for j in nb):
            na += 1
            for j in nb + [i]:
                agg[j] = na
    snap = list(agg)
    for i in range(n):
        if agg[i] is not None:
            continue
        best, pick = 0.0, None
        for j, v in strong(i):
            if snap[j] is not None and abs(v) > best:
                best, pick = abs(v), j
        if pick is not None:
            agg[i] = agg[pick]
    for i in range(n):
        if agg[i] is not None:
            continue
        na += 1
        agg[i] = na
        for j, _ in strong(i):
            if agg[j] is None:
                agg[j] = na
    return np.array(agg) - 1, na


@pytest.mark.parametrize("idim,theta", [(2, 0.0), (3, 0.0), (4, 0.0), (5, 0.0), (4, 0.2)])
def test_aggregation_two_restatements(orc, idim, theta):
    a = orc.gen_pde3d(idim)
    il, na = orc.dec_map_bld(a, theta=theta)
    il2, na2 = py_dec_map_bld(a.to_scipy(), theta)
    assert na == na2 and (il == il2).all()


def test_aggregation_golden_matrix_and_invariants(orc):
    A = orc.read_mtx(os.path.join(GOLD, "mld_mat.mtx"))
    a = orc.Csr.from_scipy(A)
    for theta in (0.0, 0.01, 0.3):
        il, na = orc.dec_map_bld(a, theta=theta)
        il2, na2 = py_dec_map_bld(A, theta)
        assert na == na2 and (il == il2).all()
        # a partition with contiguous ids 0..na-1
        assert il.min() == 0 and il.max() == na - 1 and len(np.unique(il)) == na


def test_aggregation_step1_plus_stencils(orc):
    # theta=0 on the 7-point stencil: node 0 starts aggregate 0 = {0, 1, idim, idim^2}
    idim = 6
    a = orc.gen_pde3d(idim)
    il, na = orc.dec_map_bld(a, theta=0.0)
    assert set(np.flatnonzero(il == 0)) >= {0, 1, idim, idim * idim}


# ------------------------------------------------------ smoothed prolongator
@pytest.mark.parametrize("idim", [3, 5])
def test_galerkin_and_prolongator(orc, idim):
    a = orc.gen_pde3d(idim)
    A = a.to_scipy().toarray()
    p = orc.Prec(orc.config1_params(min_coarse_size=1, max_levs=2)).build(a)
    assert p.nlev == 2
    P = p.level_csr(1, "P").to_scipy().toarray()
    R = p.level_csr(1, "R").to_scipy().toarray()
    Ac = p.level_csr(1, "A").to_scipy().toarray()
    assert (R == P.T).all()
    ref = P.T @ A @ P
    assert np.abs(Ac - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(Ac, Ac.T, rtol=0, atol=1e-12 * np.abs(Ac).max())
    # omega = 4/(3*||D^-1 A||_inf) = 2/3 on the 7-point Poisson (interior row sum 2)
    assert abs(p.omega(1) - 2.0 / 3.0) <= 4e-16
    # P = (I - omega D^-1 A) Ptilde with Ptilde from the aggregation map
    il = p.ilaggr(1)
    Pt = np.zeros((A.shape[0], p.naggr(1)))
    Pt[np.arange(A.shape[0]), il] = 1.0
    D = np.diag(1.0 / np.diag(A))
    P2 = (np.eye(A.shape[0]) - p.omega(1) * D @ A) @ Pt
    assert np.abs(P - P2).max() <= 1e-14


def test_default_coarse_size_truncates_cube_root(orc):
    # mld_d_hierarchy_bld.f90:187-197: int((N)**(1/3))*40 with the
    # floating-point cube root (64000**(1/3) = 39.99999999999999 -> 39)
    for idim, expect in [(10, 9 * 40), (20, 19 * 40), (40, 39 * 40)]:
        p = orc.Prec(orc.config1_params(min_coarse_size=-1)).build(orc.gen_pde3d(idim))
        assert p.casize == max(int(float(idim ** 3) ** (1.0 / 3.0)), 1) * 40 == expect


def test_hierarchy_stops_at_min_coarse_size(orc):
    p = orc.Prec(orc.config1_params()).build(orc.gen_pde3d(40))
    sizes = [p.level_csr(l, "A").shape[0] for l in range(p.nlev)]
    assert sizes[-1] <= 200 and all(s > 200 for s in sizes[:-1])
    assert sizes == [64000, 7809, 200]


# ------------------------------------------------------------------ ILU(0)
def ilu_check(orc, A):
    a = orc.Csr.from_scipy(A)
    L, U, d, rc = orc.ilu0(a)
    assert rc == 0
    n = A.shape[0]
    Ld = L.to_scipy().toarray() + np.eye(n)
    Ud = U.to_scipy().toarray() + np.eye(n)
    M = Ld @ np.diag(1.0 / d) @ Ud
    Ad = A.toarray()
    mask = Ad != 0
    return np.abs((M - Ad)[mask]).max() / np.abs(Ad).max(), L, U, d


def test_ilu0_defining_property(orc):
    for A in (orc.gen_pde3d(4).to_scipy(), orc.read_mtx(os.path.join(GOLD, "mld_mat.mtx"))[:400, :400].tocsr()):
        err, *_ = ilu_check(orc, A)
        assert err <= 1e-14


def test_ilu0_exact_on_tridiagonal(orc):
    n = 50
    rng = np.random.default_rng(1)
    A = sp.diags([-rng.random(n - 1), 4 + rng.random(n), -rng.random(n - 1)], [-1, 0, 1]).tocsr()
    err, L, U, d = ilu_check(orc, A)
    x = rng.standard_normal(n)
    t = orc.trsv(L, x, lower=True)
    y = orc.trsv(U, d * t, lower=False)
    assert np.allclose(A @ y, x, rtol=1e-12, atol=1e-12)
    # transposed solves give A^T y = x
    t = orc.trsv(U, x, lower=False, trans=True)
    y = orc.trsv(L, d * t, lower=True, trans=True)
    assert np.allclose(A.T @ y, x, rtol=1e-12, atol=1e-12)


def poisson2d(nx):
    T = sp.diags([-np.ones(nx - 1), 4 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    S = sp.diags([-np.ones(nx - 1), -np.ones(nx - 1)], [-1, 1])
    A = sp.kron(sp.identity(nx), T) + sp.kron(S, sp.identity(nx))
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def factors(L, U, d):
    n = len(d)
    return L.to_scipy().toarray() + np.eye(n), np.diag(1.0 / d), U.to_scipy().toarray() + np.eye(n)


def test_iluk_level_zero_is_ilu0(orc):
    """ILU(k) with k = 0 and ILU(0) are separate routines in the reference;
    on ILU(0)'s pattern they perform the same operations in the same order."""
    for A in (orc.gen_pde3d(5).to_scipy(), orc.read_mtx(os.path.join(GOLD, "mld_mat.mtx"))[:500, :500].tocsr()):
        a = orc.Csr.from_scipy(A)
        L0, U0, d0, rc0 = orc.ilu0(a)
        Lk, Uk, dk, rck = orc.iluk(a, 0)
        assert rc0 == rck == 0 and (d0 == dk).all()
        for X, Y in ((L0, Lk), (U0, Uk)):
            x, y = X.to_scipy(), Y.to_scipy()
            assert (x.indptr == y.indptr).all() and (x.indices == y.indices).all() and (x.data == y.data).all()


def test_iluk_full_fill_is_lu(orc):
    """With a fill level above any path length nothing is dropped: L D^-1 U = A
    (the exact LU of a nonsymmetric sparse matrix) and the solves give A^-1 x."""
    rng = np.random.default_rng(41)
    A = (poisson2d(6) + sp.diags([0.3 * rng.random(35)], [1], shape=(36, 36))).tocsr()
    A.sort_indices()
    L, U, d, rc = orc.iluk(orc.Csr.from_scipy(A), 40)
    assert rc == 0
    Lf, Df, Uf = factors(L, U, d)
    assert np.abs(Lf @ Df @ Uf - A.toarray()).max() <= 1e-13 * np.abs(A).max()
    x = rng.standard_normal(36)
    y = orc.trsv(U, d * orc.trsv(L, x, lower=True), lower=False)
    assert np.allclose(A @ y, x, rtol=1e-12, atol=1e-12)


def test_iluk_level_one_pattern(orc):
    """ILU(1) of the 5-point Laplacian in natural order adds exactly the fill
    diagonals at offsets -(nx-1) and nx-1 (Saad, Iterative Methods, sec. 10.3.3)."""
    nx = 7
    A = poisson2d(nx)
    L, U, d, rc = orc.iluk(orc.Csr.from_scipy(A), 1)
    assert rc == 0
    n = nx * nx
    pat = set(zip(*sp.tril(A, -1).nonzero())) | set(zip(*sp.triu(A, 1).nonzero()))
    want = pat | {(i, i - nx + 1) for i in range(n) if i - nx + 1 >= 0 and (i % nx) != nx - 1} \
               | {(i, i + nx - 1) for i in range(n) if i + nx - 1 < n and (i % nx) != 0}
    got = set(zip(*L.to_scipy().nonzero())) | set(zip(*U.to_scipy().nonzero()))
    assert got == want


@pytest.mark.parametrize("fill", [0, 1])
def test_milu_preserves_row_sums(orc, fill):
    """MILU(k) adds every dropped entry to the pivot, so the factorization
    keeps A's row sums: (I+L) D^-1 (I+U) 1 = A 1."""
    A = orc.gen_pde3d(5).to_scipy()
    L, U, d, rc = orc.iluk(orc.Csr.from_scipy(A), fill, milu=True)
    assert rc == 0
    Lf, Df, Uf = factors(L, U, d)
    e = np.ones(A.shape[0])
    assert np.allclose(Lf @ (Df @ (Uf @ e)), A @ e, rtol=1e-12, atol=1e-9 * np.abs(A).max())
    L0, U0, d0, _ = orc.iluk(orc.Csr.from_scipy(A), fill)
    Lg, Dg, Ug = factors(L0, U0, d0)
    assert np.abs(Lg @ (Dg @ (Ug @ e)) - A @ e).max() > 1e-6 * np.abs(A).max()  # plain ILU does not


# ------------------------------------------------------------- smoothers
def test_jacobi_closed_forms(orc):
    a = orc.gen_pde3d(5)
    A = a.to_scipy()
    p = orc.Prec(orc.config1_params(min_coarse_size=1, max_levs=2)).build(a)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(A.shape[0])
    d = A.diagonal()
    assert (p.smooth(0, x, sweeps=1) == (1.0 / d) * x).all()
    y1 = x / d
    y2 = (x - (A @ y1 - d * y1)) / d
    assert np.allclose(p.smooth(0, x, sweeps=2), y2, rtol=1e-14, atol=0)
    y0 = rng.standard_normal(A.shape[0])
    assert np.allclose(p.smooth(0, x, beta=1.0, y=y0, sweeps=2), y0 + y2, rtol=1e-14, atol=1e-16)
    # many sweeps converge to A^-1 x (Jacobi converges on the Poisson matrix)
    y = p.smooth(0, x, sweeps=600)
    assert np.allclose(y, spla.spsolve(A.tocsc(), x), rtol=1e-8)


def test_l1_jacobi_converges_and_diagonal(orc):
    a = orc.gen_pde3d(5)
    A = a.to_scipy()
    p = orc.Prec(orc.config1_params(smoother=orc.SM_L1JAC, min_coarse_size=1, max_levs=2)).build(a)
    d_l1 = A.diagonal() + (abs(A).sum(axis=1).A1 - abs(A.diagonal()))
    assert np.allclose(p.dv(0), 1.0 / d_l1, rtol=1e-15)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(A.shape[0])
    # l1-Jacobi: y_{k+1} = y_k + D_l1^-1 (x - A y_k), from y_0 = 0
    yk = np.zeros_like(x)
    for _ in range(3):
        yk = yk + (x - A @ yk) / d_l1
    assert np.allclose(p.smooth(0, x, sweeps=3), yk, rtol=1e-12)
    y = p.smooth(0, x, sweeps=1500)
    assert np.allclose(y, spla.spsolve(A.tocsc(), x), rtol=1e-7)


# --------------------------------------------------------------- V-cycle
def test_two_level_exact_coarse_identity(orc):
    a = orc.gen_pde3d(4)
    A = a.to_scipy().toarray()
    prm = orc.config1_params(min_coarse_size=1, max_levs=2, sweeps_pre=0, sweeps_post=0,
                             c_smoother=orc.SM_DENSELU, c_sweeps=1)
    p = orc.Prec(prm).build(a)
    P = p.level_csr(1, "P").to_scipy().toarray()
    M = P @ np.linalg.solve(P.T @ A @ P, P.T)
    x = np.random.default_rng(4).standard_normal(A.shape[0])
    assert np.allclose(p.apply(x), M @ x, rtol=1e-11, atol=1e-14)


def test_two_level_error_propagation(orc):
    # V(1,1) with Jacobi and an exact coarse solve:
    # I - M A = (I - D^-1 A)(I - P Ac^-1 P^T A)(I - D^-1 A)
    a = orc.gen_pde3d(4)
    A = a.to_scipy().toarray()
    prm = orc.config1_params(min_coarse_size=1, max_levs=2, c_smoother=orc.SM_DENSELU, c_sweeps=1)
    p = orc.Prec(prm).build(a)
    P = p.level_csr(1, "P").to_scipy().toarray()
    n = A.shape[0]
    S = np.eye(n) - np.diag(1.0 / np.diag(A)) @ A
    C = np.eye(n) - P @ np.linalg.solve(P.T @ A @ P, P.T @ A)
    E = S @ C @ S
    M = (np.eye(n) - E) @ np.linalg.inv(A)
    x = np.random.default_rng(5).standard_normal(n)
    assert np.allclose(p.apply(x), M @ x, rtol=1e-9, atol=1e-14 * np.abs(M @ x).max())


def _dense_apply(p, n, trans="N"):
    return np.column_stack([p.apply(e, trans=trans) for e in np.eye(n)])


def test_three_level_w_cycle_error_propagation(orc):
    """W-cycle (mld_dmlprec_aply.f90:716-745) on three levels, Jacobi 1/1,
    exact coarsest solve.  With Galerkin coarse operators two coarse
    corrections through a level whose own cycle has error propagator E_c are
    I - P (I - E_c^2) A_c^-1 R A, so
      E_1 = S_1 (I - P_2 A_2^-1 R_2 A_1) S_1,
      I - M A = S_0 (I - P_1 (I - E_1^2) A_1^-1 R_1 A_0) S_0,
    with S_l = I - D_l^-1 A_l; a V-cycle would have I - E_1 in place of
    I - E_1^2.  Checked on M formed column by column from the oracle."""
    a = orc.gen_pde3d(7)
    prm = orc.config1_params(min_coarse_size=8, c_smoother=orc.SM_DENSELU, c_sweeps=1, ml_cycle=orc.CYC_W)
    p = orc.Prec(prm).build(a)
    assert p.nlev == 3
    A = [p.level_csr(l, "A").to_scipy().toarray() for l in range(3)]
    P = [None] + [p.level_csr(l, "P").to_scipy().toarray() for l in (1, 2)]
    R = [None] + [p.level_csr(l, "R").to_scipy().toarray() for l in (1, 2)]
    S = [np.eye(len(Al)) - np.diag(1.0 / np.diag(Al)) @ Al for Al in A]
    n1 = len(A[1])
    E1 = S[1] @ (np.eye(n1) - P[2] @ np.linalg.solve(A[2], R[2] @ A[1])) @ S[1]
    n = len(A[0])
    E0 = S[0] @ (np.eye(n) - P[1] @ (np.eye(n1) - E1 @ E1) @ np.linalg.solve(A[1], R[1] @ A[0])) @ S[0]
    M = _dense_apply(p, n)
    Mref = (np.eye(n) - E0) @ np.linalg.inv(A[0])
    assert np.abs(M - Mref).max() <= 1e-10 * np.abs(Mref).max()
    # the V-cycle of the same hierarchy differs (the W branch is taken)
    pv = orc.Prec(orc.config1_params(min_coarse_size=8, c_smoother=orc.SM_DENSELU, c_sweeps=1)).build(a)
    assert np.abs(_dense_apply(pv, n) - Mref).max() > 1e-6 * np.abs(Mref).max()
    # symmetric A, symmetric smoothing: the W-cycle is symmetric and 'T' = 'N'
    assert np.abs(M - M.T).max() <= 1e-10 * np.abs(M).max()
    assert np.abs(_dense_apply(p, n, "T") - M).max() <= 1e-10 * np.abs(M).max()


@pytest.mark.parametrize("k", [2, 4])
def test_outer_sweeps_closed_form(orc, k):
    """outer_sweeps = k (mld_dmlprec_aply.f90:285-327): k cycles, each on the
    residual of the previous sum, so I - M_k A = (I - M A)^k, i.e.
    M_k = (I - (I - M A)^k) A^-1 with M the single-sweep preconditioner."""
    a = orc.gen_pde3d(6)
    A = a.to_scipy().toarray()
    n = len(A)
    p1 = orc.Prec(orc.config1_params(min_coarse_size=20)).build(a)
    pk = orc.Prec(orc.config1_params(min_coarse_size=20, outer_sweeps=k)).build(a)
    M = _dense_apply(p1, n)
    E = np.linalg.matrix_power(np.eye(n) - M @ A, k)
    Mk = (np.eye(n) - E) @ np.linalg.inv(A)
    x = np.random.default_rng(k).standard_normal(n)
    y = pk.apply(x)
    assert np.abs(y - Mk @ x).max() <= 1e-10 * np.abs(Mk @ x).max()
    assert np.abs(y - M @ x).max() > 1e-6 * np.abs(y).max()


def test_filtered_prolongator_dense(orc):
    """AGGR_FILTER = FILTER (mld_daggrmat_smth_asb.f90:194-217, 258-291):
    P = (I - omega D^-1 A_F) Ptilde, where A_F drops every entry with
    |a_ij| < theta sqrt(|a_ii a_jj|) and subtracts their sum from its
    diagonal, D = diag(A) and omega = 4 / (3 ||D^-1 A||_inf) come from the
    unfiltered A; Ptilde from the oracle's (separately pinned) aggregation.
    An unsmoothed heterogeneous diffusion field makes many couplings weak."""
    from mld2p4_amd import problems

    irp, ja, val, _ = problems.diffusion3d(6, seed=11, passes=0)
    a = orc.Csr.from_arrays(irp, ja, val)
    A = a.to_scipy().toarray()
    n = len(A)
    theta = 0.1
    prm = orc.config1_params(theta=theta, min_coarse_size=10, max_levs=2, aggr_filter=1)
    p = orc.Prec(prm).build(a)
    d = np.diag(A).copy()
    weak = (np.abs(A) < theta * np.sqrt(np.abs(np.outer(d, d)))) & (A != 0)
    assert weak.sum() > n // 4  # the filter removes something
    AF = np.where(weak, 0.0, A)
    AF[np.arange(n), np.arange(n)] -= np.where(weak, A, 0.0).sum(axis=1)
    Dinv = np.diag(1.0 / d)
    omega = 4.0 / (3.0 * np.abs(Dinv @ A).sum(axis=1).max())
    ilaggr = p.ilaggr(1)
    Pt = np.zeros((n, ilaggr.max() + 1))
    Pt[np.arange(n), ilaggr] = 1.0
    Pref = (np.eye(n) - omega * Dinv @ AF) @ Pt
    P = p.level_csr(1, "P").to_scipy().toarray()
    assert abs(p.omega(1) - omega) <= 1e-14 * omega
    assert np.abs(P - Pref).max() <= 1e-13 * np.abs(Pref).max()
    # unfiltered: A itself in place of A_F, and the two differ
    pu = orc.Prec(orc.config1_params(theta=theta, min_coarse_size=10, max_levs=2)).build(a)
    Pu = pu.level_csr(1, "P").to_scipy().toarray()
    assert np.abs(Pu - (np.eye(n) - omega * Dinv @ A) @ Pt).max() <= 1e-13 * np.abs(Pu).max()
    assert np.abs(P - Pu).max() > 1e-8


@pytest.mark.parametrize("sm", ["JAC", "L1JAC"])
def test_vcycle_linear_symmetric_transpose(orc, sm):
    a = orc.gen_pde3d(12)
    code = {"JAC": orc.SM_JAC, "L1JAC": orc.SM_L1JAC}[sm]
    p = orc.Prec(orc.config1_params(smoother=code, sweeps_pre=2, sweeps_post=2)).build(a)
    rng = np.random.default_rng(6)
    u, v = rng.standard_normal((2, a.shape[0]))
    Mu, Mv = p.apply(u), p.apply(v)
    assert np.allclose(p.apply(2.0 * u - 3.0 * v), 2.0 * Mu - 3.0 * Mv, rtol=1e-12, atol=1e-12 * np.abs(Mu).max())
    assert abs(u @ Mv - v @ Mu) <= 1e-12 * abs(u @ Mv)
    assert np.allclose(p.apply(u, trans="T"), Mu, rtol=1e-11, atol=1e-12 * np.abs(Mu).max())


# --------------------------------------------------- hybrid Gauss-Seidel
def test_gauss_seidel_sweeps(orc):
    """GS: one sweep from zero is forward substitution with tril(A) (scipy's
    triangular solver); BWGS: backward substitution with triu(A); a second
    sweep is (D+L)^-1 (x - U y1) (dense solve); 'T' is refused as in the
    reference solver."""
    a = orc.gen_pde3d(5)
    A = a.to_scipy()
    prm = orc.config1_params(smoother=orc.SM_FBGS, sweeps_pre=1, sweeps_post=1, min_coarse_size=20)
    p = orc.Prec(prm).build(a)
    x = np.random.default_rng(12).standard_normal(A.shape[0])
    Lo = sp.tril(A, 0, format="csr")
    Up = sp.triu(A, 0, format="csr")
    y_gs = p.smooth(0, x, which=0, sweeps=1)
    y_bw = p.smooth(0, x, which=1, sweeps=1)
    assert np.allclose(y_gs, spla.spsolve_triangular(Lo, x, lower=True), rtol=1e-13, atol=0)
    assert np.allclose(y_bw, spla.spsolve_triangular(Up, x, lower=False), rtol=1e-13, atol=0)
    Ad = A.toarray()
    Ls, Us = np.tril(Ad), np.triu(Ad, 1)
    y1 = np.linalg.solve(Ls, x)
    y2 = np.linalg.solve(Ls, x - Us @ y1)
    assert np.allclose(p.smooth(0, x, which=0, sweeps=2), y2, rtol=1e-12, atol=1e-14 * np.abs(y2).max())
    # y = alpha S x + beta y
    y0 = np.random.default_rng(13).standard_normal(A.shape[0])
    assert np.allclose(p.smooth(0, x, alpha=1.0, beta=1.0, y=y0, which=1, sweeps=1), y_bw + y0, rtol=1e-15)
    with pytest.raises(RuntimeError):
        p.smooth(0, x, trans="T")


def test_fbgs_two_level_error_propagation_and_symmetry(orc):
    """V(1,1) with forward GS before and backward GS after an exact coarse
    solve: I - M A = (I - (D+U)^-1 A)(I - P Ac^-1 P^T A)(I - (D+L)^-1 A); the
    resulting M is symmetric for symmetric A."""
    a = orc.gen_pde3d(4)
    A = a.to_scipy().toarray()
    prm = orc.config1_params(smoother=orc.SM_FBGS, min_coarse_size=1, max_levs=2, c_smoother=orc.SM_DENSELU,
                             c_sweeps=1)
    p = orc.Prec(prm).build(a)
    P = p.level_csr(1, "P").to_scipy().toarray()
    n = A.shape[0]
    Spre = np.eye(n) - np.linalg.solve(np.tril(A), A)
    Spost = np.eye(n) - np.linalg.solve(np.triu(A), A)
    C = np.eye(n) - P @ np.linalg.solve(P.T @ A @ P, P.T @ A)
    M = (np.eye(n) - Spost @ C @ Spre) @ np.linalg.inv(A)
    x = np.random.default_rng(14).standard_normal(n)
    assert np.allclose(p.apply(x), M @ x, rtol=1e-9, atol=1e-14 * np.abs(M @ x).max())
    rng = np.random.default_rng(15)
    a3 = orc.gen_pde3d(10)
    p3 = orc.Prec(orc.config1_params(smoother=orc.SM_FBGS, sweeps_pre=2, sweeps_post=2)).build(a3)
    u, v = rng.standard_normal((2, a3.shape[0]))
    assert abs(u @ p3.apply(v) - v @ p3.apply(u)) <= 1e-12 * abs(u @ p3.apply(v))


@pytest.mark.parametrize("nsw", [1, 2])
def test_fbgs_one_level_symmetrized_pair(orc, nsw):
    """One-level FBGS (MAX_LEVS 1): nsw times w2 = (D+L)^-1 w1, w1 = (D+U)^-1 w2
    from w1 = x, each a single sweep from zero; for 'T' the two solvers
    swap, and GS refuses 'T'."""
    a = orc.gen_pde3d(5)
    A = a.to_scipy()
    prm = orc.config1_params(smoother=orc.SM_FBGS, sweeps_pre=nsw, sweeps_post=1, max_levs=1)
    p = orc.Prec(prm).build(a)
    assert p.nlev == 1
    x = np.random.default_rng(16).standard_normal(A.shape[0])
    Lo = sp.tril(A, 0, format="csr")
    Up = sp.triu(A, 0, format="csr")
    w = x.copy()
    for _ in range(nsw):
        w = spla.spsolve_triangular(Up, spla.spsolve_triangular(Lo, w, lower=True), lower=False)
    assert np.allclose(p.apply(x), w, rtol=1e-12, atol=1e-14 * np.abs(w).max())
    with pytest.raises(RuntimeError):
        p.apply(x, trans="T")


# ------------------------------------------------------- ADD and K cycles
@pytest.mark.parametrize("smoother", ["JACOBI", "FBGS"])
def test_add_cycle_closed_form(orc, smoother):
    """Additive two-level cycle with an exact coarse solve:
    M = S + P A_c^-1 R, S = D^-1 (one Jacobi sweep) or (D+U)^-1 (D+L)^-1 (the
    GS / BWGS pair of FBGS, applied in turn)."""
    a = orc.gen_pde3d(6)
    A = a.to_scipy().toarray()
    sm = orc.SM_JAC if smoother == "JACOBI" else orc.SM_FBGS
    prm = orc.config1_params(smoother=sm, min_coarse_size=1, max_levs=2, c_smoother=orc.SM_DENSELU, c_sweeps=1,
                             ml_cycle=orc.CYC_ADD)
    p = orc.Prec(prm).build(a)
    assert p.nlev == 2
    P = p.level_csr(1, "P").to_scipy().toarray()
    R = p.level_csr(1, "R").to_scipy().toarray()
    Ac = p.level_csr(1, "A").to_scipy().toarray()
    if smoother == "JACOBI":
        S = np.diag(1.0 / np.diag(A))
    else:
        S = np.linalg.solve(np.triu(A), np.linalg.inv(np.tril(A)))
    M = S + P @ np.linalg.solve(Ac, R)
    x = np.random.default_rng(31).standard_normal(A.shape[0])
    assert np.allclose(p.apply(x), M @ x, rtol=1e-10, atol=1e-13 * np.abs(M @ x).max())


def test_kcycle_two_levels_is_vcycle(orc):
    """With two levels the K-cycle has no inner iteration: pre-smoothing,
    coarse correction, post-smoothing, i.e. the V-cycle (and MULT is the
    V-cycle), bit for bit."""
    a = orc.gen_pde3d(8)
    x = np.random.default_rng(32).standard_normal(a.shape[0])
    ys = []
    for cyc in (orc.CYC_V, orc.CYC_MULT, orc.CYC_K, orc.CYC_KSYM):
        prm = orc.config1_params(smoother=orc.SM_JAC, sweeps_pre=2, sweeps_post=2, min_coarse_size=60, max_levs=2,
                                 ml_cycle=cyc)
        p = orc.Prec(prm).build(a)
        assert p.nlev == 2
        ys.append(p.apply(x))
    assert all((y == ys[0]).all() for y in ys[1:])


@pytest.mark.parametrize("cyc", ["K", "KSYM"])
@pytest.mark.parametrize("case", ["one_step", "two_steps"])
def test_kcycle_inner_iteration_is_krylov_projection(orc, cyc, case):
    """Three levels: the level-1 correction of the K-cycle is the K-cycle
    iterate of at most two steps preconditioned by B_1 (the level-1 cycle),
    here formed independently as a projection onto span{d0, d1}, d0 = B_1 r,
    d1 = B_1 (r - alpha A_1 d0): the A_1-orthogonal (Galerkin) projection for
    the symmetric variant (FCG), the least-squares residual minimiser for
    GCR; one step (x = alpha d0) when ||r - alpha A_1 d0|| <= ||r|| / 4.
    A weak B_1 (no smoother) forces the second step."""
    a = orc.gen_pde3d(10 if case == "one_step" else 12)
    ml = orc.CYC_K if cyc == "K" else orc.CYC_KSYM
    sm = orc.SM_JAC if case == "one_step" else orc.SM_NOPREC
    prm = orc.config1_params(smoother=sm, min_coarse_size=20 if case == "one_step" else 60, ml_cycle=ml)
    p = orc.Prec(prm).build(a)
    assert p.nlev == 3
    A0 = a.to_scipy()
    A1 = p.level_csr(1, "A").to_scipy()
    P1 = p.level_csr(1, "P").to_scipy()
    R1 = p.level_csr(1, "R").to_scipy()
    B1 = np.column_stack([p.level_cycle(1, e) for e in np.eye(A1.shape[0])])
    rng = np.random.default_rng(33)
    for _ in range(3):
        x = rng.standard_normal(A0.shape[0])
        y = p.smooth(0, x)
        r = R1 @ (x - A0 @ y)
        d0 = B1 @ r
        v = A1 @ d0
        alpha = (d0 @ r) / (d0 @ v) if cyc == "KSYM" else (v @ r) / (v @ v)
        w = r - alpha * v
        one = np.linalg.norm(w) <= 0.25 * np.linalg.norm(r)
        assert one == (case == "one_step")
        if one:
            y1 = alpha * d0
        else:
            D = np.column_stack([d0, B1 @ w])
            if cyc == "KSYM":
                c = np.linalg.solve(D.T @ (A1 @ D), D.T @ r)
            else:
                c = np.linalg.lstsq(A1 @ D, r, rcond=None)[0]
            y1 = D @ c
        y = y + P1 @ y1
        y = p.smooth(0, x - A0 @ y, beta=1.0, y=y)
        yo = p.apply(x)
        assert np.abs(yo - y).max() <= 1e-12 * np.abs(yo).max()


# ----------------------------------------------------------------- Krylov
@pytest.mark.parametrize("method", ["CG", "FCG", "BICGSTAB"])
def test_krylov_small_exact(orc, method):
    a = orc.gen_pde3d(3)
    A = a.to_scipy()
    p = orc.Prec(orc.config1_params(min_coarse_size=5)).build(a)
    b = np.random.default_rng(7).standard_normal(A.shape[0])
    x, it, err = p.krylov(method, b, eps=1e-13, itmax=A.shape[0] + 5)
    assert err <= 1e-13 and it <= A.shape[0]
    assert np.allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-10)


@pytest.mark.parametrize("method", ["CG", "FCG", "BICGSTAB"])
def test_golden_solution(orc, method):
    """tests/fileread/runs/mld_{mat,rhs,sol}.mtx of the reference: solving to
    1e-12 reproduces the shipped solution to kappa*1e-10."""
    A = orc.read_mtx(os.path.join(GOLD, "mld_mat.mtx"))
    b = orc.read_mtx(os.path.join(GOLD, "mld_rhs.mtx"))
    xr = orc.read_mtx(os.path.join(GOLD, "mld_sol.mtx"))
    p = orc.Prec(orc.config1_params()).build(orc.Csr.from_scipy(A))
    x, it, err = p.krylov(method, b, eps=1e-12, itmax=500)
    assert err <= 1e-12
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 2e-7
