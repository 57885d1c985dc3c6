"""GPU parity of the engine's device path against the oracle, through the C-ABI.

Tolerances: FP64 preconditioned vector <= 1e-12 relative (max-norm), the
bar stated for the FP64 apply; single sparse products whose kernel keeps the
stored per-row order (ELL, one lane per row) are compared bit-for-bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


def engine_prec(eng, A, smoother="JACOBI", sweeps=1, coarse="BJAC", csweeps=4, mcs=200, precision="double",
                fmt="auto", extra=()):
    p = eng.MldPrec(device=0).init("ML")
    p.set("SMOOTHER_TYPE", smoother)
    p.set("SMOOTHER_SWEEPS", sweeps)
    p.set("AGGR_THRESH", 0.0)
    p.set("AGGR_PROL", "SMOOTHED")
    p.set("MIN_COARSE_SIZE", mcs)
    p.set("COARSE_SOLVE", coarse)
    if coarse == "BJAC":
        p.set("COARSE_SUBSOLVE", "ILU")
    p.set("COARSE_SWEEPS", csweeps)
    for k, v in extra:
        p.set(k, v)
    p.set_precision(precision)
    p.set_format(fmt)
    return p.build(A)


SM = {"NOPREC": 0, "JACOBI": 1, "BJAC": 2, "L1-JACOBI": 100, "FBGS": 4}
CYC = {"ADD": 1, "MULT": 2, "VCYCLE": 3, "WCYCLE": 4, "KCYCLE": 5, "KCYCLESYM": 6}


def oracle_prec(orc, a, smoother="JACOBI", sweeps=1, coarse="BJAC", csweeps=4, mcs=200, **kw):
    cs = {"BJAC": orc.SM_BJAC, "DENSE_LU": orc.SM_DENSELU, "JACOBI": orc.SM_JAC}[coarse]
    prm = orc.config1_params(smoother=SM[smoother], sweeps_pre=sweeps, sweeps_post=sweeps, c_smoother=cs,
                             c_sweeps=csweeps, min_coarse_size=mcs, **kw)
    return orc.Prec(prm).build(a)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def relmax(y, yo):
    return float(np.abs(y - yo).max() / max(np.abs(yo).max(), 1e-300))


@pytest.mark.parametrize("idim", [6, 12, 24])
@pytest.mark.parametrize("smoother,sweeps", [("JACOBI", 1), ("JACOBI", 2), ("L1-JACOBI", 2), ("L1-JACOBI", 3),
                                             ("BJAC", 1), ("FBGS", 1), ("FBGS", 2)])
def test_apply_fp64(eng, orc, idim, smoother, sweeps):
    _say(f"[apply fp64] idim={idim} smoother={smoother} sweeps={sweeps}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, smoother, sweeps)
    po = oracle_prec(orc, a, smoother, sweeps)
    assert p.nlev == po.nlev
    rng = np.random.default_rng(idim)
    for trial in range(2):
        x = rng.standard_normal(A.shape[0])
        y = p.apply(dev(x)).cpu().numpy()
        yo = po.apply(x)
        assert relmax(y, yo) <= 1e-12


@pytest.mark.parametrize("fmt", ["auto", "csr", "ell", "sell", "dia", "seld", "seld4096", "seld128"])
def test_level_spmv_bitexact(eng, orc, fmt):
    _say(f"[level spmv] fmt={fmt}")
    a = orc.gen_pde3d(20)
    A = a.to_scipy()
    p = engine_prec(eng, A, fmt=fmt)
    po = oracle_prec(orc, a)
    rng = np.random.default_rng(3)
    for lev in range(p.nlev):
        for what in (["A", "ND"] if lev == 0 else ["A", "P", "R", "ND"]):
            if what == "ND" and p.level_info(lev)["nnz_nd"] == 0:
                continue
            m = po.level_csr(lev, what)
            nr, nc = m.shape
            x = rng.standard_normal(nc)
            y = torch.zeros(nr, dtype=torch.float64, device="cuda:0")
            p.level_spmv(lev, what, dev(x), y)
            yo = m.spmv(x)
            yg = y.cpu().numpy()
            f = p.op_format(lev, what)
            if fmt.startswith("seld"):
                assert f["fmt"] == "SELD", (lev, what, f)
            if f["fmt"] in ("ELL", "SELL", "DIA", "SELD") or f["lanes"] == 1:
                # one lane per row, stored order, no fused multiply-add:
                # identical to the sequential CSR loop
                assert (yg == yo).all(), (lev, what, f)
            else:
                assert relmax(yg, yo) <= 1e-14, (lev, what, f)


@pytest.mark.parametrize("smoother,sweeps,beta", [("JACOBI", 1, 0), ("JACOBI", 2, 0), ("JACOBI", 3, 1),
                                                  ("L1-JACOBI", 2, 1), ("L1-JACOBI", 4, 0), ("BJAC", 1, 0),
                                                  ("BJAC", 1, 1)])
def test_level_smoother(eng, orc, smoother, sweeps, beta):
    _say(f"[smoother] {smoother} sweeps={sweeps} beta={beta}")
    a = orc.gen_pde3d(14)
    A = a.to_scipy()
    p = engine_prec(eng, A, smoother, sweeps)
    po = oracle_prec(orc, a, smoother, sweeps)
    rng = np.random.default_rng(5)
    n = A.shape[0]
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    y = dev(y0.copy())
    p.level_smooth(0, dev(x), y, beta=beta, sweeps=sweeps)
    yo = po.smooth(0, x, alpha=1.0, beta=float(beta), y=y0, sweeps=sweeps)
    assert relmax(y.cpu().numpy(), yo) <= 1e-13


@pytest.mark.parametrize("idim,beta", [(30, 0), (30, 1), (52, 0)])
def test_ilu_level_scheduled_bitexact(eng, orc, idim, beta):
    """ILU(0) smoother on levels large enough for the sync-free substitution
    (rows in level-schedule order taken by workgroups through a ticket, each
    product waiting for the value it reads; 4 or 16 lanes per row) and on a
    small coarse level (one workgroup, level sets with barriers): identical
    to the sequential substitution, for 'N' and for 'T' (column-oriented
    order), on repeated applications."""
    _say(f"[ILU level-scheduled] idim={idim} beta={beta}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, "BJAC", 1)
    po = oracle_prec(orc, a, "BJAC", 1)
    rng = np.random.default_rng(8)
    for lev in range(min(2, p.nlev - 1)):
        n = p.level_info(lev)["n"]
        x = rng.standard_normal(n)
        y0 = rng.standard_normal(n)
        for trans in ("N", "T", "N"):
            y = dev(y0.copy())
            p.level_smooth(lev, dev(x), y, beta=beta, sweeps=1, trans=trans)
            yo = po.smooth(lev, x, alpha=1.0, beta=float(beta), y=y0, sweeps=1, trans=trans)
            assert (y.cpu().numpy() == yo).all(), (lev, trans)


@pytest.mark.parametrize("idim", [30, 41, 52])
def test_ilu_grid_substitution_bitexact(eng, orc, idim, capfd, monkeypatch):
    """ILU(0) factors of the 7-point stencil on the fine level are recognised
    as a grid (the upload says so under MLD_VERBOSE) and solved by
    trsv_grid_kernel (tiles of 8 x 8 grid lines; idim 30, 41, 52 leave
    partial tiles): 'N' applications with beta 0 and 1 identical to the
    oracle's sequential substitution, and to the general sync-free form
    (MLD_TRSV_GRID=0) in FP64 and in FP32."""
    _say(f"[ILU grid] idim={idim}")
    monkeypatch.setenv("MLD_VERBOSE", "1")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, "BJAC", 1)
    out = capfd.readouterr().err
    assert f"on a {idim} x {idim} x {idim} grid: grid substitutions" in out
    po = oracle_prec(orc, a, "BJAC", 1)
    rng = np.random.default_rng(17)
    n = idim ** 3
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    for beta in (0, 1, 0):
        monkeypatch.delenv("MLD_TRSV_GRID", raising=False)
        y = dev(y0.copy())
        p.level_smooth(0, dev(x), y, beta=beta, sweeps=1, trans="N")
        yo = po.smooth(0, x, alpha=1.0, beta=float(beta), y=y0, sweeps=1, trans="N")
        assert (y.cpu().numpy() == yo).all(), beta
        monkeypatch.setenv("MLD_TRSV_GRID", "0")
        y2 = dev(y0.copy())
        p.level_smooth(0, dev(x), y2, beta=beta, sweeps=1, trans="N")
        assert torch.equal(y, y2), beta
    monkeypatch.delenv("MLD_TRSV_GRID", raising=False)
    p32 = engine_prec(eng, A, "BJAC", 1, precision="mixed")
    x32 = dev(x)
    z1 = p32.apply(x32)
    monkeypatch.setenv("MLD_TRSV_GRID", "0")
    z2 = p32.apply(x32)
    assert torch.equal(z1, z2)


@pytest.mark.parametrize("idim", [14, 30])
def test_gauss_seidel_level_bitexact(eng, orc, idim):
    """GS (forward) and BWGS (backward) smoothers of the default ML
    preconditioner, 1 and 3 sweeps, beta 0 and 1: level-scheduled
    substitutions (one workgroup for idim 14, grid launches and runs for
    idim 30) identical to the oracle's sequential sweeps; 'T' is refused."""
    _say(f"[GS level] idim={idim}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, "FBGS", 1)
    po = oracle_prec(orc, a, "FBGS", 1)
    rng = np.random.default_rng(21)
    n = A.shape[0]
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    for which in (0, 1):
        for sweeps in (1, 3):
            for beta in (0, 1):
                y = dev(y0.copy())
                p.level_smooth(0, dev(x), y, beta=beta, sweeps=sweeps, which=which)
                yo = po.smooth(0, x, alpha=1.0, beta=float(beta), y=y0, sweeps=sweeps, which=which)
                assert (y.cpu().numpy() == yo).all(), (which, sweeps, beta)
    with pytest.raises(eng.MldError):
        p.level_smooth(0, dev(x), dev(y0.copy()), trans="T")


@pytest.mark.parametrize("idim", [41])
def test_gauss_seidel_grid_substitution_bitexact(eng, orc, idim, capfd, monkeypatch):
    """GS and BWGS sweeps on the fine level run as grid substitutions
    (trsv_grid_kernel with the division by the stored diagonal): identical to
    the oracle's sequential sweeps and to the general sync-free form
    (MLD_TRSV_GRID=0), 1 and 3 sweeps, beta 0 and 1."""
    _say(f"[GS grid] idim={idim}")
    monkeypatch.setenv("MLD_VERBOSE", "1")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, "FBGS", 1)
    out = capfd.readouterr().err
    assert f"Gauss-Seidel factors of a 7-point stencil on a {idim} x {idim} x {idim} grid" in out
    po = oracle_prec(orc, a, "FBGS", 1)
    rng = np.random.default_rng(23)
    n = A.shape[0]
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    for which in (0, 1):
        for sweeps in (1, 3):
            for beta in (0, 1):
                monkeypatch.delenv("MLD_TRSV_GRID", raising=False)
                y = dev(y0.copy())
                p.level_smooth(0, dev(x), y, beta=beta, sweeps=sweeps, which=which)
                yo = po.smooth(0, x, alpha=1.0, beta=float(beta), y=y0, sweeps=sweeps, which=which)
                assert (y.cpu().numpy() == yo).all(), (which, sweeps, beta)
                monkeypatch.setenv("MLD_TRSV_GRID", "0")
                y2 = dev(y0.copy())
                p.level_smooth(0, dev(x), y2, beta=beta, sweeps=sweeps, which=which)
                assert torch.equal(y, y2), (which, sweeps, beta)


@pytest.mark.parametrize("nsw", [1, 2])
def test_one_level_fbgs_apply(eng, orc, nsw):
    """init('FBGS'): the symmetrized one-level pair, bit-identical to the
    oracle's; 'T' is refused as by the GS solver."""
    _say(f"[one-level FBGS] nsw={nsw}")
    a = orc.gen_pde3d(24)
    A = a.to_scipy()
    p = eng.MldPrec(device=0).init("FBGS")
    p.set("SMOOTHER_SWEEPS", nsw)
    p.build(A)
    po = orc.Prec(orc.config1_params(smoother=orc.SM_FBGS, sweeps_pre=nsw, sweeps_post=nsw, max_levs=1)).build(a)
    x = np.random.default_rng(22).standard_normal(A.shape[0])
    y = p.apply(dev(x)).cpu().numpy()
    assert (y == po.apply(x)).all()
    with pytest.raises(eng.MldError):
        p.apply(dev(x), trans="T")


@pytest.mark.parametrize("solve,fill", [("ILU", 2), ("MILU", 1)])
def test_iluk_smoother_bitexact(eng, orc, solve, fill):
    """ILU(k) / MILU(k) factors in the level-scheduled solves ('N' and 'T'):
    bit-identical to the oracle's sequential substitution; the whole
    preconditioner within the FP64 bar."""
    _say(f"[ILU(k)] {solve}({fill}) idim=24")
    a = orc.gen_pde3d(24)
    A = a.to_scipy()
    p = engine_prec(eng, A, "BJAC", 1, extra=(("SUB_SOLVE", solve), ("SUB_FILLIN", fill)))
    po = oracle_prec(orc, a, "BJAC", 1, sub_fill=fill, sub_milu=int(solve == "MILU"))
    rng = np.random.default_rng(24)
    n = A.shape[0]
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    for trans in ("N", "T"):
        y = dev(y0.copy())
        p.level_smooth(0, dev(x), y, beta=1, sweeps=1, trans=trans)
        yo = po.smooth(0, x, alpha=1.0, beta=1.0, y=y0, sweeps=1, trans=trans)
        assert (y.cpu().numpy() == yo).all(), trans
    assert relmax(p.apply(dev(x)).cpu().numpy(), po.apply(x)) <= 1e-12


def test_additive_schwarz_apply(eng, orc):
    """AS on one process: the multilevel apply equals BJAC's bit for bit and
    the oracle's within the FP64 bar; one-level AS with the diagonal solver
    applies D^-1 once whatever the sweeps."""
    _say("[AS] idim=20")
    a = orc.gen_pde3d(20)
    A = a.to_scipy()
    p_as = engine_prec(eng, A, "AS", 1)
    p_bj = engine_prec(eng, A, "BJAC", 1)
    po = orc.Prec(orc.config1_params(smoother=orc.SM_AS)).build(a)
    x = np.random.default_rng(23).standard_normal(A.shape[0])
    y = p_as.apply(dev(x)).cpu().numpy()
    assert (y == p_bj.apply(dev(x)).cpu().numpy()).all()
    assert relmax(y, po.apply(x)) <= 1e-12
    q = eng.MldPrec(device=0).init("AS")
    q.set("SUB_SOLVE", "DIAG")
    q.set("SMOOTHER_SWEEPS", 3)
    q.build(A)
    assert (q.apply(dev(x)).cpu().numpy() == x * (1.0 / A.diagonal())).all()


def test_krylov_default_ml(eng, orc):
    """The reference's default ML preconditioner (FBGS smoother, coarse ILU(0)
    through block Jacobi) in CG: the same iteration count as the oracle."""
    _say("[krylov default ML] idim=30")
    a = orc.gen_pde3d(30)
    A = a.to_scipy()
    p = engine_prec(eng, A, "FBGS", 1)
    po = oracle_prec(orc, a, "FBGS", 1)
    b = np.ones(A.shape[0])
    x, it, err = eng.psb_krylov("CG", A, p, dev(b), eps=1e-8)
    xo, ito, erro = po.krylov("CG", b, eps=1e-8)
    assert err <= 1e-8 and abs(it - ito) <= 1
    r = b - A @ x.cpu().numpy()
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-8


@pytest.mark.parametrize("cycle,smoother,sweeps,idim,mcs", [
    ("ADD", "JACOBI", 1, 16, 40), ("ADD", "FBGS", 1, 16, 40), ("MULT", "L1-JACOBI", 2, 16, 40),
    ("WCYCLE", "JACOBI", 1, 16, 40), ("WCYCLE", "L1-JACOBI", 2, 24, 40), ("WCYCLE", "FBGS", 1, 20, 40),
    ("WCYCLE", "BJAC", 1, 20, 40),
    ("KCYCLE", "JACOBI", 2, 24, 40), ("KCYCLESYM", "JACOBI", 2, 24, 40),
    ("KCYCLE", "NOPREC", 1, 12, 60), ("KCYCLESYM", "NOPREC", 1, 12, 60), ("KCYCLESYM", "FBGS", 1, 30, 40)])
def test_cycles_vs_oracle(eng, orc, cycle, smoother, sweeps, idim, mcs):
    """ADD, MULT and K-cycles (GCR and FCG inner iterations; the NOPREC
    smoother makes the inner iteration take its second step) against the
    oracle.  The K-cycle's inner dots are reduced in blocks on the device and
    sequentially in the oracle, so agreement is to rounding, not bit-exact."""
    _say(f"[cycle] {cycle} {smoother} sweeps={sweeps} idim={idim}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    p = engine_prec(eng, A, smoother, sweeps, mcs=mcs, extra=(("ML_CYCLE", cycle),))
    po = oracle_prec(orc, a, smoother, sweeps, mcs=mcs, ml_cycle=CYC[cycle])
    assert p.nlev == po.nlev
    if cycle.startswith("K"):
        assert p.nlev >= 3
    rng = np.random.default_rng(idim + 7)
    for trial in range(2):
        x = rng.standard_normal(A.shape[0])
        y = p.apply(dev(x)).cpu().numpy()
        assert relmax(y, po.apply(x)) <= 1e-11, trial
    if smoother in ("JACOBI", "NOPREC"):
        x = rng.standard_normal(A.shape[0])
        y = p.apply(dev(x), trans="T").cpu().numpy()
        assert relmax(y, po.apply(x, trans="T")) <= 1e-11


@pytest.mark.parametrize("k,precision", [(2, "double"), (4, "double"), (3, "mixed")])
def test_outer_sweeps_vs_oracle(eng, orc, k, precision):
    """OUTER_SWEEPS = k (mld_dmlprec_aply.f90:285-327: k cycles on the
    residual of the running sum): within 1e-12 of the oracle's application
    in FP64 (1e-5 with the FP32 hierarchy), 'N' and 'T'; CG converges."""
    _say(f"[outer sweeps] k={k} {precision} idim=16")
    a = orc.gen_pde3d(16)
    A = a.to_scipy()
    p = engine_prec(eng, A, "L1-JACOBI", 2, mcs=40, precision=precision, extra=(("OUTER_SWEEPS", k),))
    po = oracle_prec(orc, a, "L1-JACOBI", 2, mcs=40, outer_sweeps=k)
    x = np.random.default_rng(40 + k).standard_normal(A.shape[0])
    tol = 1e-12 if precision == "double" else 1e-5
    for trans in ("N", "T"):
        assert relmax(p.apply(dev(x), trans=trans).cpu().numpy(), po.apply(x, trans=trans)) <= tol, trans
    b = np.ones(A.shape[0])
    xs, it, err = eng.psb_krylov("CG", A, p, dev(b), eps=1e-8)
    assert err <= 1e-8
    assert np.linalg.norm(b - A @ xs.cpu().numpy()) / np.linalg.norm(b) <= 1.01e-8


def test_kcycle_mixed_and_krylov(eng, orc):
    """FCG with the symmetric K-cycle (the reference pairs them): FP64 and
    FP32-mixed hierarchies reach the tolerance in the oracle's iteration
    count (+-1)."""
    _say("[K-cycle krylov] idim=32")
    a = orc.gen_pde3d(32)
    A = a.to_scipy()
    po = oracle_prec(orc, a, "L1-JACOBI", 1, mcs=40, ml_cycle=CYC["KCYCLESYM"])
    b = np.ones(A.shape[0])
    xo, ito, erro = po.krylov("FCG", b, eps=1e-8)
    for prec in ("double", "mixed"):
        p = engine_prec(eng, A, "L1-JACOBI", 1, mcs=40, precision=prec, extra=(("ML_CYCLE", "KCYCLESYM"),))
        x, it, err = eng.psb_krylov("FCG", A, p, dev(b), eps=1e-8)
        assert err <= 1e-8 and abs(it - ito) <= 1, (prec, it, ito)
        r = b - A @ x.cpu().numpy()
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-8


@pytest.mark.parametrize("method,smoother,sweeps", [("CG", "L1-JACOBI", 2), ("FCG", "L1-JACOBI", 2),
                                                   ("FCG", "JACOBI", 1), ("CG", "BJAC", 1), ("FCG", "BJAC", 1)])
def test_mixed_krylov_fp32_output(eng, orc, method, smoother, sweeps):
    """FP32-mixed CG/FCG keep the preconditioner output in FP32 (fused into
    the last Jacobi sweep with its dot, or left by the ILU / single-sweep
    paths and dotted separately): the oracle's FP64 iteration count (+-1) and
    the true residual."""
    _say(f"[mixed krylov] {method} {smoother} sweeps={sweeps} idim=28")
    a = orc.gen_pde3d(28)
    A = a.to_scipy()
    po = oracle_prec(orc, a, smoother, sweeps)
    b = np.ones(A.shape[0])
    _, ito, _ = po.krylov(method, b, eps=1e-8)
    p = engine_prec(eng, A, smoother, sweeps, precision="mixed")
    x, it, err = eng.psb_krylov(method, A, p, dev(b), eps=1e-8)
    assert err <= 1e-8 and abs(it - ito) <= 1, (it, ito)
    r = b - A @ x.cpu().numpy()
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-8


@pytest.mark.parametrize("method", ["CG", "FCG", "BICGSTAB"])
def test_krylov_config1(eng, orc, method):
    _say(f"[krylov] {method} idim=40")
    a = orc.gen_pde3d(40)
    A = a.to_scipy()
    p = engine_prec(eng, A)
    po = oracle_prec(orc, a)
    b = np.ones(A.shape[0])
    x, it, err = eng.psb_krylov(method, A, p, dev(b), eps=1e-6)
    xo, ito, erro = po.krylov(method, b, eps=1e-6)
    assert err <= 1e-6
    assert abs(it - ito) <= 1
    r = b - A @ x.cpu().numpy()
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-6


def test_apply_mixed_fp32(eng, orc):
    _say("[apply mixed] idim=24")
    a = orc.gen_pde3d(24)
    A = a.to_scipy()
    p = engine_prec(eng, A, "L1-JACOBI", 2, precision="mixed")
    po = oracle_prec(orc, a, "L1-JACOBI", 2)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(A.shape[0])
    y = p.apply(dev(x)).cpu().numpy()
    yo = po.apply(x)
    assert relmax(y, yo) <= 1e-5


def test_apply_transpose_symmetric(eng, orc):
    _say("[apply T] idim=12")
    a = orc.gen_pde3d(12)
    A = a.to_scipy()
    p = engine_prec(eng, A, "JACOBI", 2)
    po = oracle_prec(orc, a, "JACOBI", 2)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(A.shape[0])
    yt = p.apply(dev(x), trans="T").cpu().numpy()
    yto = po.apply(x, trans="T")
    assert relmax(yt, yto) <= 1e-12


def test_dense_lu_coarse_apply(eng, orc):
    """DENSE_LU coarsest solve: device blocked LU (MFMA updates) + inverse."""
    _say("[dense LU coarse] idim=14")
    a = orc.gen_pde3d(14)
    A = a.to_scipy()
    p = engine_prec(eng, A, "JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=600)
    po = oracle_prec(orc, a, "JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=600)
    assert p.nlev == po.nlev and p.level_info(p.nlev - 1)["solver"] == 102
    assert p.level_info(p.nlev - 1)["n"] > 100
    rng = np.random.default_rng(9)
    x = rng.standard_normal(A.shape[0])
    assert relmax(p.apply(dev(x)).cpu().numpy(), po.apply(x)) <= 1e-12


def test_dense_lu_two_level_identity(eng, orc):
    """No smoothing and an exact coarse solve: M = P (P^T A P)^-1 P^T."""
    _say("[dense LU identity] idim=10")
    a = orc.gen_pde3d(10)
    A = a.to_scipy()
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "JACOBI"), ("SMOOTHER_SWEEPS", 0), ("AGGR_THRESH", 0.0), ("MAX_LEVS", 2),
                 ("MIN_COARSE_SIZE", 5), ("COARSE_SOLVE", "DENSE_LU")]:
        p.set(k, v)
    p.build(A)
    P = p.level_csr(1, "P").toarray()
    Ad = A.toarray()
    M = P @ np.linalg.solve(P.T @ Ad @ P, P.T)
    x = np.random.default_rng(4).standard_normal(A.shape[0])
    y = p.apply(dev(x)).cpu().numpy()
    assert relmax(y, M @ x) <= 1e-10


def test_diffusion_mixed_fcg(eng, orc):
    """Heterogeneous diffusion, FP32 hierarchy inside FP64 FCG, dense-LU coarse:
    converges to 1e-8 and matches the oracle's FP64 solution."""
    _say("[diffusion mixed] idim=32")
    from mld2p4_amd import problems

    irp, ja, val, b = problems.diffusion3d(32, seed=3)
    a = orc.Csr.from_arrays(irp, ja, val)
    A = a.to_scipy()
    p = engine_prec(eng, A, "L1-JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=4096, precision="mixed")
    x, it, err = eng.psb_krylov("FCG", A, p, dev(b), eps=1e-8)
    assert err <= 1e-8
    po = oracle_prec(orc, a, "L1-JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=4096)
    xo, ito, erro = po.krylov("CG", b, eps=1e-12, itmax=300)
    xg = x.cpu().numpy()
    rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"mixed FCG iters {it} (oracle FP64 CG to 1e-12: {ito}); solution error vs oracle {rel:.2e}")
    assert rel <= 1e-7


def test_host_vector_entry_points(eng, orc):
    """mldhip_apply_host / mldhip_krylov_host (the Fortran driver's path):
    copies through pinned staging into device vectors the handle keeps.  One
    handle is built for idim 24 and then rebuilt for idim 164 (the kept
    vectors grow: freed and allocated again on the same handle; a vector is
    35 MB, two chunks of the 32 MB staging buffers).  After each build the
    host-vector results equal the device-vector ones bit for bit on repeated
    calls, the application equals the oracle's within 1e-12, and CG reaches
    1e-8 with the true residual checked on the host."""
    _say("[host vectors] one handle, idim 24 then 164")
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                 ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"),
                 ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    for idim in (24, 164):
        a = orc.gen_pde3d(idim)
        A = a.to_scipy()
        p.hierarchy_build(A)
        p.smoothers_build()
        po = oracle_prec(orc, a, "L1-JACOBI", 2)
        n = A.shape[0]
        x = np.random.default_rng(idim).standard_normal(n)
        for _ in range(2):
            yh = p.apply(x)
            yd = p.apply(dev(x)).cpu().numpy()
            assert (yh == yd).all()
        assert relmax(yh, po.apply(x)) <= 1e-12
        b = np.ones(n)
        xh, ith, erh = eng.psb_krylov("CG", A, p, b, eps=1e-8)
        xd, itd, erd = eng.psb_krylov("CG", A, p, dev(b), eps=1e-8)
        assert ith == itd and erh == erd
        assert (xh == xd.cpu().numpy()).all()
        assert erh <= 1e-8
        assert np.linalg.norm(b - A @ xh) / np.linalg.norm(b) <= 1.01e-8


def test_apply_after_error_return(eng, orc):
    """An application that returns an error (the dense-LU coarsest solver
    refuses 'T') leaves the handle usable: the next 'N' application equals the
    oracle's within 1e-12, also with the FP32 hierarchy in FCG."""
    _say("[apply after error] idim=14")
    a = orc.gen_pde3d(14)
    A = a.to_scipy()
    po = oracle_prec(orc, a, "JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=600)
    x = np.random.default_rng(31).standard_normal(A.shape[0])
    yo = po.apply(x)
    for prec, tol in (("double", 1e-12), ("mixed", 1e-5)):
        p = engine_prec(eng, A, "JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=600, precision=prec)
        with pytest.raises(eng.MldError):
            p.apply(dev(x), trans="T")
        assert relmax(p.apply(dev(x)).cpu().numpy(), yo) <= tol
        b = np.ones(A.shape[0])
        _, it, err = eng.psb_krylov("FCG", A, p, dev(b), eps=1e-8)
        assert err <= 1e-8


@pytest.mark.parametrize("cycle,smoother,sweeps,coarse,idim", [
    ("VCYCLE", "L1-JACOBI", 2, "BJAC", 16), ("WCYCLE", "JACOBI", 1, "BJAC", 16), ("WCYCLE", "L1-JACOBI", 2, "BJAC", 24),
    ("VCYCLE", "BJAC", 1, "BJAC", 24), ("VCYCLE", "FBGS", 1, "BJAC", 16), ("ADD", "JACOBI", 1, "BJAC", 16),
    ("VCYCLE", "L1-JACOBI", 2, "DENSE_LU", 24), ("KCYCLE", "JACOBI", 2, "BJAC", 24)])
def test_dense_tail(eng, orc, monkeypatch, cycle, smoother, sweeps, coarse, idim):
    """The cycle on the levels from the first one of at most 1024 rows below
    the fine level, applied as one dense product (formed at build by running
    that cycle on the unit vectors) in 'N' applications: the same operator as
    the level-by-level cycle of the same hierarchy (MLD_DENSE_TAIL=0) and
    within 1e-11 of the oracle, FP64 and FP32; 'T' (where the solvers allow
    it) keeps the cycle; CG iterations equal.  Not taken with a K-cycle below
    that level."""
    _say(f"[dense tail] {cycle} {smoother} sweeps={sweeps} coarse={coarse} idim={idim}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    csw = 1 if coarse == "DENSE_LU" else 4
    kw = dict(mcs=40, coarse=coarse, csweeps=csw, extra=(("ML_CYCLE", cycle),))
    p = engine_prec(eng, A, smoother, sweeps, **kw)
    monkeypatch.setenv("MLD_DENSE_TAIL", "0")
    q = engine_prec(eng, A, smoother, sweeps, **kw)
    q32 = engine_prec(eng, A, smoother, sweeps, precision="mixed", **kw)
    monkeypatch.delenv("MLD_DENSE_TAIL")
    p32 = engine_prec(eng, A, smoother, sweeps, precision="mixed", **kw)
    assert q.dense_tail() == -1 and q32.dense_tail() == -1
    sizes = [p.level_info(l)["n"] for l in range(p.nlev)]
    expect = next((l for l in range(1, p.nlev) if sizes[l] <= 1024), -1)
    if cycle.startswith("K") or (expect == p.nlev - 1 and coarse == "DENSE_LU"):
        expect = -1
    assert p.dense_tail() == expect and p32.dense_tail() == expect, sizes
    assert expect >= 1 or cycle.startswith("K")
    po = oracle_prec(orc, a, smoother, sweeps, mcs=40, coarse=coarse, csweeps=csw, ml_cycle=CYC[cycle])
    rng = np.random.default_rng(idim + 11)
    for trans in (("N", "T") if smoother != "FBGS" and coarse != "DENSE_LU" else ("N",)):
        x = rng.standard_normal(A.shape[0])
        yo = po.apply(x, trans=trans)
        yp = p.apply(dev(x), trans=trans).cpu().numpy()
        yq = q.apply(dev(x), trans=trans).cpu().numpy()
        assert relmax(yp, yq) <= 1e-12 and relmax(yp, yo) <= 1e-11, trans
        y32 = p32.apply(dev(x), trans=trans).cpu().numpy()
        assert relmax(y32, q32.apply(dev(x), trans=trans).cpu().numpy()) <= 1e-5
        assert relmax(y32, yo) <= 1e-5, trans
    b = np.ones(A.shape[0])
    _, itp, errp = eng.psb_krylov("CG", A, p, dev(b), eps=1e-8)
    _, itq, _ = eng.psb_krylov("CG", A, q, dev(b), eps=1e-8)
    assert errp <= 1e-8 and abs(itp - itq) <= 1
