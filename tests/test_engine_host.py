"""CPU tests of the engine library through its C-ABI (no GPU): exported
symbols, the configuration API's routing, the host hierarchy build and
smoother data against the oracle (bit-exact: integers and values), error
codes, and the problem generator / byte model used by the benchmark."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_capi_exports_every_declared_symbol(eng):
    hdr = open(os.path.join(ROOT, "include", "mld2p4_amd.h")).read()
    names = set(re.findall(r"\bint\s+(mldhip_\w+)\s*\(", hdr))
    assert len(names) >= 30
    lib = eng.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def config1(eng, smoother="JACOBI", sweeps=1, coarse="BJAC", csweeps=4, mcs=200):
    p = eng.MldPrec(device=-1).init("ML")
    p.set("SMOOTHER_TYPE", smoother)
    p.set("SMOOTHER_SWEEPS", sweeps)
    p.set("AGGR_THRESH", 0.0)
    p.set("AGGR_PROL", "SMOOTHED")
    p.set("MIN_COARSE_SIZE", mcs)
    p.set("COARSE_SOLVE", coarse)
    if coarse == "BJAC":
        p.set("COARSE_SUBSOLVE", "ILU")
    p.set("COARSE_SWEEPS", csweeps)
    return p


def same(m1, m2):
    return (m1.shape == m2.shape and (m1.indptr == m2.indptr).all() and (m1.indices == m2.indices).all()
            and (m1.data == m2.data).all())


SMC = {"JACOBI": 1, "L1-JACOBI": 100, "BJAC": 2, "FBGS": 4}


@pytest.mark.parametrize("idim", [7, 20, 40])
@pytest.mark.parametrize("smoother", ["JACOBI", "L1-JACOBI", "BJAC", "FBGS"])
def test_host_build_bitexact_vs_oracle(eng, orc, idim, smoother):
    a = orc.gen_pde3d(idim)
    p = config1(eng, smoother, 2).build(a.to_scipy())
    po = orc.Prec(orc.config1_params(smoother=SMC[smoother], sweeps_pre=2, sweeps_post=2)).build(a)
    assert p.nlev == po.nlev
    for l in range(p.nlev):
        assert same(p.level_csr(l, "A"), po.level_csr(l, "A").to_scipy())
        if l > 0:
            assert (p.aggr(l) == po.ilaggr(l)).all()
            assert same(p.level_csr(l, "P"), po.level_csr(l, "P").to_scipy())
            assert same(p.level_csr(l, "R"), po.level_csr(l, "R").to_scipy())
            assert p.omega(l) == po.omega(l)
        info = p.level_info(l)
        if info["nnz_nd"] > 0:
            assert same(p.level_csr(l, "ND"), po.level_csr(l, "ND").to_scipy())
        assert (p.dv(l) == po.dv(l)).all()
        if info["solver"] in (6, 7, 14):  # GS, ILU(0), BWGS
            assert same(p.level_csr(l, "L"), po.level_csr(l, "L").to_scipy())
            assert same(p.level_csr(l, "U"), po.level_csr(l, "U").to_scipy())
        if smoother == "FBGS" and l < p.nlev - 1:
            assert (info["solver"], info["solver_post"]) == (6, 14)
            assert (p.dv(l, which=1) == po.dv(l)).all()


def test_host_build_filtered_prolongator_vs_oracle(eng, orc):
    """AGGR_FILTER = FILTER on a heterogeneous diffusion matrix whose weak
    couplings the filter drops: aggregation, P, R and A_c equal the oracle's
    bit for bit."""
    from mld2p4_amd import problems

    irp, ja, val, _ = problems.diffusion3d(10, seed=11, passes=0)
    a = orc.Csr.from_arrays(irp, ja, val)
    p = config1(eng, "JACOBI", 1, mcs=20)
    p.set("AGGR_THRESH", 0.1)
    p.set("AGGR_FILTER", "FILTER")
    p.build((irp, ja, val))
    po = orc.Prec(orc.config1_params(theta=0.1, min_coarse_size=20, aggr_filter=1)).build(a)
    pu = orc.Prec(orc.config1_params(theta=0.1, min_coarse_size=20)).build(a)
    assert p.nlev == po.nlev >= 3
    for l in range(1, p.nlev):
        assert (p.aggr(l) == po.ilaggr(l)).all()
        for w in ("P", "R", "A"):
            assert same(p.level_csr(l, w), po.level_csr(l, w).to_scipy()), (l, w)
        assert p.omega(l) == po.omega(l)
    assert not same(po.level_csr(1, "P").to_scipy(), pu.level_csr(1, "P").to_scipy())


def test_one_level_fbgs_build(eng, orc):
    """init('FBGS'): one level, GS before and BWGS after, the split of A as the
    oracle's one-level build (MAX_LEVS 1) makes it."""
    a = orc.gen_pde3d(9)
    p = eng.MldPrec(device=-1).init("FBGS").build(a.to_scipy())
    po = orc.Prec(orc.config1_params(smoother=orc.SM_FBGS, max_levs=1)).build(a)
    assert p.nlev == po.nlev == 1
    info = p.level_info(0)
    assert (info["solver"], info["solver_post"]) == (6, 14)
    for w in ("L", "U"):
        assert same(p.level_csr(0, w), po.level_csr(0, w).to_scipy())
    assert (p.dv(0) == po.dv(0)).all()


def test_host_build_golden_matrix(eng, orc):
    A = orc.read_mtx(os.path.join(GOLD, "mld_mat.mtx"))
    for theta in (0.0, 0.05):
        p = config1(eng)
        p.set("AGGR_THRESH", theta)
        p.build(A)
        po = orc.Prec(orc.config1_params(theta=theta)).build(orc.Csr.from_scipy(A))
        assert p.nlev == po.nlev
        for l in range(1, p.nlev):
            assert (p.aggr(l) == po.ilaggr(l)).all()
            assert same(p.level_csr(l, "A"), po.level_csr(l, "A").to_scipy())


def test_config_routing(eng):
    p = eng.MldPrec(device=-1).init("ML")
    # ML defaults (mld_dprecinit.F90:188-207): FBGS smoother, coarse ILU
    p.set("SMOOTHER_TYPE", "L1-JACOBI")
    p.set("SMOOTHER_SWEEPS", 3, pos="post")
    p.set("COARSE_SOLVE", "BJAC")
    p.set("COARSE_SWEEPS", 4)
    p.set("MIN_COARSE_SIZE", 50)
    from mld2p4_amd import problems

    irp, ja, val = problems.pde3d_poisson(12)
    p.hierarchy_build((irp, ja, val)).smoothers_build()
    nl = p.nlev
    assert nl >= 2
    for l in range(nl - 1):
        i = p.level_info(l)
        assert (i["solver"], i["sweeps_pre"], i["sweeps_post"]) == (101, 1, 3)
    c = p.level_info(nl - 1)
    assert (c["solver"], c["sweeps_pre"], c["sweeps_post"]) == (7, 4, 4)
    assert c["n"] <= 50


def test_error_codes(eng):
    E = eng.MldError
    p = eng.MldPrec(device=-1)
    with pytest.raises(E) as ex:
        p.hierarchy_build(((0, 1), (0,), (1.0,)))
    assert ex.value.info == 3111
    with pytest.raises(E) as ex:
        p.init("NOT-A-PREC")
    assert ex.value.info == 2
    p.init("ML")
    p.set("SMOOTHER_TYPE", "BJAC")
    p.set("SUB_SOLVE", "ILUT")  # ILU with threshold: not provided by this engine
    from mld2p4_amd import problems

    A = problems.pde3d_poisson(6)
    p.hierarchy_build(A)
    with pytest.raises(E) as ex:
        p.smoothers_build()
    assert ex.value.info == -20
    # unsorted CSR rows are rejected
    q = eng.MldPrec(device=-1).init("JACOBI")
    with pytest.raises(E) as ex:
        q.hierarchy_build((np.array([0, 2, 3]), np.array([1, 0, 1]), np.array([1.0, 2.0, 3.0])))
    assert ex.value.info == -1
    # apply before build
    r = eng.MldPrec(device=-1).init("JACOBI")
    with pytest.raises(E) as ex:
        r.apply(np.ones(3))
    assert ex.value.info == 3112
    # zero pivot in ILU(0)
    s = eng.MldPrec(device=-1).init("BJAC")
    s.hierarchy_build((np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([1.0, 1.0, 1.0, 1.0])))
    with pytest.raises(E) as ex:
        s.smoothers_build()
    assert ex.value.info == 2
    # built on a host-only handle: no device copy
    t = eng.MldPrec(device=-1).init("JACOBI")
    t.build(A).descr()
    with pytest.raises(E) as ex:
        t.apply(np.ones(216))
    assert ex.value.info == -30


def test_one_level_jacobi_and_zero_diagonal(eng):
    # mld_d_diag_solver_bld.f90:86-92: a zero diagonal is replaced by 1
    p = eng.MldPrec(device=-1).init("JACOBI")
    p.build((np.array([0, 2, 3]), np.array([0, 1, 0]), np.array([4.0, 1.0, 2.0])))
    assert (p.dv(0) == np.array([0.25, 1.0])).all()


def test_problem_generator_matches_oracle(eng, orc):
    from mld2p4_amd import problems

    for idim in (3, 11):
        irp, ja, val = problems.pde3d_poisson(idim)
        io, jo, vo, _ = orc.gen_pde3d(idim).arrays()
        assert (irp == io).all() and (ja == jo).all() and (val == vo).all()


def test_byte_model(eng):
    from mld2p4_amd import perfmodel

    # fine-level residual at 256^3 in CSR: values, column indices, row pointers, three vectors
    n, nnz = 256 ** 3, 117047296
    assert perfmodel.residual_bytes(n, nnz, 8, 8, csr=True) == 1874329604
    # two fused l1-Jacobi sweeps from a zero guess: ND + x + d + y
    assert perfmodel.smoother_bytes(10, 101, 2, 70, 0, 8, 8, 0) == 70 * 12 + 10 * 24
    tot, steps = perfmodel.vcycle_bytes(
        [dict(n=10, nnz_a=70, nnz_p=0, nnz_r=0, nnz_nd=70, solver=101, sweeps_pre=2, sweeps_post=2),
         dict(n=3, nnz_a=9, nnz_p=20, nnz_r=20, nnz_nd=0, solver=7, sweeps_pre=4, sweeps_post=4, nnz_lu=6)], 8, 8)
    assert len(steps) == 7 and tot == sum(s[2] for s in steps)


def test_diffusion_generator_properties(eng):
    """Heterogeneous FV diffusion: symmetric, zero row sums on interior rows
    (flux conservation), positive boundary row sums (Dirichlet), harmonic
    face coefficients, seeded and reproducible."""
    import scipy.sparse as sp
    from mld2p4_amd import problems

    idim = 9
    irp, ja, val, b = problems.diffusion3d(idim, seed=11)
    irp2, ja2, val2, b2 = problems.diffusion3d(idim, seed=11)
    assert (val == val2).all() and (b == b2).all()
    n = idim ** 3
    A = sp.csr_matrix((val, ja, irp), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    rs = np.asarray(A.sum(axis=1)).ravel()
    g = np.arange(n)
    x, y, z = g // idim ** 2, (g // idim) % idim, g % idim
    inner = (x > 0) & (x < idim - 1) & (y > 0) & (y < idim - 1) & (z > 0) & (z < idim - 1)
    assert np.abs(rs[inner]).max() <= 1e-12 * A.diagonal().max()
    assert rs[~inner].min() > 0
    # one face by hand: k from the smoothed field, harmonic mean, / h^2
    rng = np.random.Generator(np.random.Philox(11))
    f = rng.standard_normal((idim, idim, idim))
    for _ in range(3):
        f = problems.box3(f)
    k = np.exp(2.0 * f).reshape(n)
    i = (4 * idim + 4) * idim + 4
    assert np.isclose(A[i, i + 1], -2 * k[i] * k[i + 1] / (k[i] + k[i + 1]) * idim ** 2, rtol=1e-14)
    assert np.isclose(A[i, i + idim * idim], -2 * k[i] * k[i + idim * idim] / (k[i] + k[i + idim * idim]) * idim ** 2,
                      rtol=1e-14)
    assert b.min() >= -1.0 and b.max() <= 1.0


def test_dump_matrix_market(eng, orc, tmp_path):
    """prec%dump: file names of mld_d_base_onelev_dump / the smoother and solver
    dumps, contents equal to the level data (17 significant digits: exact);
    GS writes tril(A) with the diagonal and BWGS triu(A) with the diagonal, as
    their builds store them."""
    import scipy.io
    import scipy.sparse as sp

    a = orc.gen_pde3d(12)
    A = a.to_scipy()
    p = config1(eng, "FBGS", 1).build(A)
    assert p.nlev >= 3
    prefix = str(tmp_path / "mld")
    p.dump(prefix, istart=1, ac=True, rp=True, smoother=True, solver=True, head="test dump")

    def rd(name):
        m = scipy.io.mmread(f"{prefix}{name}")
        return sp.csr_matrix(m) if sp.issparse(m) else np.asarray(m).ravel()

    def eq(m1, m2):
        d = (sp.csr_matrix(m1) - sp.csr_matrix(m2))
        return m1.shape == m2.shape and (d.count_nonzero() == 0)

    assert not os.path.exists(f"{prefix}_p000_l001_ac.mtx")  # level 1 holds no ac
    for lev in range(2, p.nlev + 1):
        l = lev - 1
        assert eq(rd(f"_p000_l{lev:03d}_ac.mtx"), p.level_csr(l, "A"))
        assert eq(rd(f"_p000_l{lev:03d}_r.mtx"), p.level_csr(l, "R"))
        assert eq(rd(f"_p000_l{lev:03d}_p.mtx"), p.level_csr(l, "P"))
    for lev in range(1, p.nlev):
        Al = p.level_csr(lev - 1, "A")
        assert eq(rd(f"_sm_p{0:03d}_l{lev:03d}_lower.mtx"), sp.tril(Al, 0))
        assert eq(rd(f"_sm_p000_l{lev:03d}_upper.mtx"), sp.triu(Al, 1))
        assert eq(rd(f"_sm2a_p000_l{lev:03d}_lower.mtx"), sp.tril(Al, -1))
        assert eq(rd(f"_sm2a_p000_l{lev:03d}_upper.mtx"), sp.triu(Al, 0))
        assert rd(f"_sm_p000_l{lev:03d}_nd.mtx").nnz == 0
    lc = p.nlev
    assert eq(rd(f"_sm_p000_l{lc:03d}_lower.mtx"), p.level_csr(lc - 1, "L"))
    assert eq(rd(f"_sm_p000_l{lc:03d}_upper.mtx"), p.level_csr(lc - 1, "U"))
    assert (rd(f"_sm_p000_l{lc:03d}_diag.mtx") == p.dv(lc - 1)).all()
    assert "% test dump" in open(f"{prefix}_p000_l002_ac.mtx").read(200)
    # default range: levels 2..nlev, level data only when asked
    q = str(tmp_path / "d")
    p.dump(q, ac=True)
    assert os.path.exists(f"{q}_p000_l002_ac.mtx") and not os.path.exists(f"{q}_p000_l002_r.mtx")


def test_additive_schwarz_one_process(eng, orc):
    """AS on one process is the block-Jacobi smoother with an empty ND
    (no overlap rows, identity restriction/prolongation): same ILU(0)
    factors as BJAC; with SUB_SOLVE DIAG only the inverse diagonal, no ND
    (unlike the Jacobi smoother).  SUB_SOLVE JACOBI names no solver of the
    reference's list (mld_d_base_onelev_seti.F90:148-192) and keeps ILU(0)."""
    a = orc.gen_pde3d(10)
    A = a.to_scipy()
    p_as = config1(eng, "AS", 1).build(A)
    p_bj = config1(eng, "BJAC", 1).build(A)
    po = orc.Prec(orc.config1_params(smoother=orc.SM_AS)).build(a)
    assert p_as.nlev == p_bj.nlev == po.nlev
    for l in range(p_as.nlev):
        ia, ib = p_as.level_info(l), p_bj.level_info(l)
        assert ia["solver"] == ib["solver"] == 7 and ia["nnz_nd"] == 0
        for w in ("L", "U"):
            assert same(p_as.level_csr(l, w), p_bj.level_csr(l, w))
            assert same(p_as.level_csr(l, w), po.level_csr(l, w).to_scipy())
        assert (p_as.dv(l) == p_bj.dv(l)).all()
    q = eng.MldPrec(device=-1).init("AS")
    q.set("SUB_SOLVE", "DIAG")
    q.build(A)
    info = q.level_info(0)
    assert info["nnz_nd"] == 0 and info["solver"] == 5
    assert (q.dv(0) == 1.0 / A.diagonal()).all()
    r = eng.MldPrec(device=-1).init("AS")
    r.set("SUB_SOLVE", "JACOBI")
    assert r.build(A).level_info(0)["solver"] == 7


@pytest.mark.parametrize("solve,fill", [("ILU", 1), ("ILU", 2), ("MILU", 0), ("MILU", 1)])
def test_iluk_milu_factors_vs_oracle(eng, orc, solve, fill):
    """BJAC smoother with ILU(k) or MILU(k) (SUB_SOLVE / SUB_FILLIN): factors
    and inverse pivots bit-identical to the oracle's."""
    a = orc.gen_pde3d(12)
    p = config1(eng, "BJAC", 1)
    p.set("SUB_SOLVE", solve)
    p.set("SUB_FILLIN", fill)
    p.build(a.to_scipy())
    po = orc.Prec(orc.config1_params(smoother=orc.SM_BJAC, sub_fill=fill, sub_milu=int(solve == "MILU"))).build(a)
    assert p.nlev == po.nlev
    for l in range(p.nlev):
        for w in ("L", "U"):
            assert same(p.level_csr(l, w), po.level_csr(l, w).to_scipy()), (l, w)
        assert (p.dv(l) == po.dv(l)).all()
    if fill > 0:
        assert p.level_csr(0, "L").nnz > sp_tril_nnz(a.to_scipy())


def sp_tril_nnz(A):
    import scipy.sparse as sp

    return sp.tril(A, -1).nnz
