"""A hierarchy built outside the engine (CPU, no GPU): levels and smoother
data handed over through mldhip_upload_level / mldhip_upload_smoother /
mldhip_finalize_hierarchy, or read back from a prec%dump by
mldhip_load_hierarchy, give the same host data, bit for bit, as the engine's
own hierarchy_build + smoothers_build; malformed input is refused with the
documented codes."""
import numpy as np
import pytest

SOLVER_NAME = {4: "NONE", 5: "DIAG", 6: "GS", 7: "ILU", 14: "BWGS", 101: "L1-DIAG"}


def config(eng, smoother, sweeps=2, device=-1):
    p = eng.MldPrec(device=device).init("ML")
    for k, v in [("SMOOTHER_TYPE", smoother), ("SMOOTHER_SWEEPS", sweeps), ("AGGR_THRESH", 0.0),
                 ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", 60), ("COARSE_SOLVE", "BJAC"),
                 ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    return p


def same(m1, m2):
    return (m1.shape == m2.shape and (m1.indptr == m2.indptr).all() and (m1.indices == m2.indices).all()
            and (m1.data == m2.data).all())


def assert_same_prec(p, q):
    assert p.nlev == q.nlev
    for l in range(p.nlev):
        ip, iq = p.level_info(l), q.level_info(l)
        for k in ("n", "nnz_a", "nnz_p", "nnz_r", "nnz_nd", "sweeps_pre", "sweeps_post", "solver", "solver_post"):
            assert ip[k] == iq[k], (l, k, ip[k], iq[k])
        ops = ["A", "ND"] + (["P", "R"] if l > 0 else [])
        if ip["solver"] in (6, 7, 14):
            ops += ["L", "U"]
        if ip["solver_post"] in (6, 7, 14):
            ops += ["L2", "U2"]
        for w in ops:
            assert same(p.level_csr(l, w), q.level_csr(l, w)), (l, w)
        assert (p.dv(l) == q.dv(l)).all()
        if ip["solver_post"] >= 0:
            assert (p.dv(l, which=1) == q.dv(l, which=1)).all()


def export_upload(p, q, smoothers):
    for l in range(p.nlev):
        if l == 0:
            q.upload_level(0, p.level_csr(0, "A"))
        else:
            q.upload_level(l, p.level_csr(l, "A"), p.level_csr(l, "P"), p.level_csr(l, "R"))
        if smoothers:
            info = p.level_info(l)
            for which, code, (wl, wu, wnd) in ((0, info["solver"], ("L", "U", "ND")),
                                              (1, info["solver_post"], ("L2", "U2", "ND2"))):
                if code < 0:
                    continue
                tri = code in (6, 7, 14)
                nd = p.level_csr(l, wnd) if which == 0 and info["nnz_nd"] > 0 else None
                q.upload_smoother(l, SOLVER_NAME[code], dv=p.dv(l, which=which), nd=nd,
                                  l=p.level_csr(l, wl) if tri else None, u=p.level_csr(l, wu) if tri else None,
                                  which=which)
    q.finalize_hierarchy(p.nlev)


@pytest.mark.parametrize("smoother", ["JACOBI", "L1-JACOBI", "BJAC", "FBGS"])
@pytest.mark.parametrize("smoothers", [False, True])
def test_upload_equals_built(eng, orc, smoother, smoothers):
    A = orc.gen_pde3d(14).to_scipy()
    p = config(eng, smoother).build(A)
    assert p.nlev >= 3
    q = config(eng, smoother)
    export_upload(p, q, smoothers)
    if not smoothers:
        q.smoothers_build()
    assert_same_prec(p, q)


@pytest.mark.parametrize("smoother", ["L1-JACOBI", "BJAC", "FBGS"])
@pytest.mark.parametrize("smoothers", [False, True])
def test_load_dump_equals_built(eng, orc, tmp_path, smoother, smoothers):
    A = orc.gen_pde3d(13).to_scipy()
    p = config(eng, smoother).build(A)
    prefix = str(tmp_path / "h")
    p.dump(prefix, istart=1, ac=True, rp=True, smoother=True, solver=True)
    q = config(eng, smoother).load_hierarchy(prefix, A, smoothers=smoothers)
    if not smoothers:
        q.smoothers_build()
    assert_same_prec(p, q)


def test_upload_refusals(eng, orc):
    A = orc.gen_pde3d(8).to_scipy()
    p = config(eng, "JACOBI").build(A)
    q = config(eng, "JACOBI")
    # a level's smoother before the level itself
    with pytest.raises(eng.MldError) as e:
        q.upload_smoother(0, "DIAG", dv=p.dv(0))
    assert e.value.info == -2
    # columns not ascending in a row
    import scipy.sparse as sp

    bad = A.tocsr().copy()
    bad.indices[[0, 1]] = bad.indices[[1, 0]]
    bad.has_sorted_indices = True
    with pytest.raises(eng.MldError) as e:
        q.upload_level(0, (bad.indptr, bad.indices, bad.data))
    assert e.value.info == -1
    # P of level 1 whose rows do not match level 0
    q.upload_level(0, p.level_csr(0, "A"))
    P = p.level_csr(1, "P")
    q.upload_level(1, p.level_csr(1, "A"), sp.csr_matrix(P[:-1]), sp.csr_matrix(p.level_csr(1, "R")[:, :-1]))
    with pytest.raises(eng.MldError) as e:
        q.finalize_hierarchy(2)
    assert e.value.info == -1
    # smoother data for some levels only
    q.upload_level(1, p.level_csr(1, "A"), P, p.level_csr(1, "R"))
    q.upload_smoother(0, "DIAG", dv=p.dv(0), nd=p.level_csr(0, "ND"))
    with pytest.raises(eng.MldError) as e:
        q.finalize_hierarchy(2)
    assert e.value.info == -1
    # a triangle holding its diagonal
    with pytest.raises(eng.MldError) as e:
        q.upload_smoother(1, "ILU", dv=np.ones(P.shape[1]), l=sp.tril(p.level_csr(1, "A")).tocsr(),
                          u=sp.triu(p.level_csr(1, "A"), 1).tocsr())
    assert e.value.info == -1
    # a solver with no data layout here
    with pytest.raises(eng.MldError) as e:
        eng.lib()
        from ctypes import c_void_p

        rc = eng.lib().mldhip_upload_smoother(q.h, 0, 0, 102, None, None, None, None, None, None, None, None, None,
                                              None)
        q._check(rc, "upload_smoother")
    assert e.value.info == -20
    # finalize without any staged level
    r = config(eng, "JACOBI")
    with pytest.raises(eng.MldError) as e:
        r.finalize_hierarchy(1)
    assert e.value.info == -1


def test_load_dump_missing_smoother_files(eng, orc, tmp_path):
    """The default dump range starts at level 2: loading smoothers from it
    names the missing level-1 file."""
    A = orc.gen_pde3d(10).to_scipy()
    p = config(eng, "JACOBI").build(A)
    prefix = str(tmp_path / "d")
    p.dump(prefix, ac=True, rp=True, smoother=True, solver=True)
    with pytest.raises(eng.MldError) as e:
        config(eng, "JACOBI").load_hierarchy(prefix, A, smoothers=True)
    assert e.value.info == -1 and "l001" in str(e.value)
    q = config(eng, "JACOBI").load_hierarchy(prefix, A).smoothers_build()
    assert_same_prec(p, q)
