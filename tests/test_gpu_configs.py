"""The benchmark workloads at their stated sizes, checked against the oracle
through the C-ABI.

Small-size parity (tests/test_gpu_parity.py) does not reach the kernel forms
the large levels use: the DIA products with non-temporal output stores (levels
of 2^22 rows and more), SELL storage of a 2 M-row second level, and the layouts
the timing autotuner picks at these sizes.  Each test here runs one workload
at the size the benchmark times:

  * 3-D Poisson idim 256 (16.8 M rows), L1-JACOBI 2/2: host hierarchy
    bit-identical to the oracle's (aggregation, P, R, A_c of every level),
    preconditioner applications within 1e-12 of the oracle's, CG to 1e-8 in the
    oracle's iteration count with the true residual checked on the host, and
    the FP32-mixed FCG solution against the oracle's FP64 solution;
  * the same problem with one BJAC/ILU(0) sweep;
  * 256^3 heterogeneous diffusion: FP32-mixed FCG to 1e-8 within 1e-7 of the
    oracle's FP64 CG solution to 1e-12 (dense-LU coarsest level on both sides);
  * 192^3 Poisson with 32 Gaussian right-hand sides (block V-cycle, block CG);
  * 64 Poisson systems with idim uniform in [24, 96] (grouped V-cycle, one
    graph for the iterations of all systems);
  * the dL/dtheta workload at idim 128 against central differences and against
    a gradient formed from the oracle's FP64 solves.

The oracle's row loops run on all host threads (its results do not depend on
the thread count).  Tolerances are stated in each test.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


SM = {"JACOBI": 1, "BJAC": 2, "L1-JACOBI": 100}


def engine_prec(eng, A, smoother, sweeps, coarse="BJAC", csweeps=4, mcs=200, precision="double"):
    p = eng.MldPrec(device=0).init("ML")
    p.set("SMOOTHER_TYPE", smoother)
    p.set("SMOOTHER_SWEEPS", sweeps)
    if smoother == "BJAC":
        p.set("SUB_SOLVE", "ILU")
    p.set("AGGR_THRESH", 0.0)
    p.set("AGGR_PROL", "SMOOTHED")
    p.set("MIN_COARSE_SIZE", mcs)
    p.set("COARSE_SOLVE", coarse)
    if coarse == "BJAC":
        p.set("COARSE_SUBSOLVE", "ILU")
    p.set("COARSE_SWEEPS", csweeps)
    p.set_precision(precision)
    return p.build(A)


def oracle_prec(orc, a, smoother, sweeps, coarse="BJAC", csweeps=4, mcs=200):
    cs = {"BJAC": orc.SM_BJAC, "DENSE_LU": orc.SM_DENSELU}[coarse]
    prm = orc.config1_params(smoother=SM[smoother], sweeps_pre=sweeps, sweeps_post=sweeps, c_smoother=cs,
                             c_sweeps=csweeps, min_coarse_size=mcs)
    return orc.Prec(prm).build(a)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def relmax(y, yo):
    return float(np.abs(y - yo).max() / max(np.abs(yo).max(), 1e-300))


def rel2(y, yo):
    return float(np.linalg.norm(y - yo) / np.linalg.norm(yo))


def true_relres(problems, irp, ja, val, x, b):
    return float(np.linalg.norm(b - problems.csr_matvec(irp, ja, val, x)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def threads(orc):
    return orc.set_threads(0)


@pytest.fixture(scope="module")
def poisson256(eng, orc):
    from mld2p4_amd import problems

    irp, ja, val = problems.pde3d_poisson(256)
    a = orc.Csr.from_arrays(irp, ja, val)
    return irp, ja, val, a


def _level_arrays_equal(p, po, lev, what):
    m = p.level_csr(lev, what)
    irp, ja, val, shape = po.level_csr(lev, what).arrays()
    assert m.shape == shape, (lev, what)
    assert np.array_equal(m.indptr.astype(np.int64), irp), (lev, what, "irp")
    assert np.array_equal(m.indices, ja), (lev, what, "ja")
    assert np.array_equal(m.data, val), (lev, what, "val")


def test_config2_poisson256_l1jacobi(eng, orc, threads, poisson256):
    """Poisson idim 256, L1-JACOBI 2/2, CG to 1e-8: hierarchy bit-identical to
    the oracle's; two applications within 1e-12 (max-norm, relative) of the
    oracle's, through the DIA non-temporal-store kernels of the fine level and
    the autotuned layouts of the coarser ones; CG in the oracle's iteration
    count (+-1) with the true residual <= 1.01e-8; the FP32-mixed FCG solution
    within 1e-7 of the oracle's FP64 CG solution."""
    from mld2p4_amd import problems

    irp, ja, val, a = poisson256
    n = len(irp) - 1
    _say(f"[config 2, run A] idim=256 n={n}, oracle threads {threads}")
    t0 = time.perf_counter()
    p = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2)
    t1 = time.perf_counter()
    po = oracle_prec(orc, a, "L1-JACOBI", 2)
    t2 = time.perf_counter()
    _say(f"  engine build {t1 - t0:.1f}s, oracle build {t2 - t1:.1f}s, nlev {p.nlev}")
    assert p.nlev == po.nlev
    # host data: aggregation maps and level operators, bit for bit
    for lev in range(1, p.nlev):
        assert np.array_equal(p.aggr(lev), po.ilaggr(lev)), lev
        for what in ("A", "P", "R"):
            _level_arrays_equal(p, po, lev, what)
    assert np.array_equal(p.dv(0), po.dv(0))
    # the fine level runs as DIA with the non-temporal-store epilogue form
    assert p.op_format(0, "A")["fmt"] == "DIA" and n >= (1 << 22)
    fmts = {lev: (p.op_format(lev, "A")["fmt"], p.op_format(lev, "A")["lanes"]) for lev in range(p.nlev)}
    _say(f"  layouts of A: {fmts}")
    rng = np.random.default_rng(256)
    for trial in range(2):
        x = rng.standard_normal(n)
        y = p.apply(dev(x)).cpu().numpy()
        yo = po.apply(x)
        r = relmax(y, yo)
        _say(f"  apply {trial}: rel diff {r:.2e}")
        assert r <= 1e-12
    b = np.ones(n)
    xg, it, err = eng.psb_krylov("CG", None, p, dev(b), eps=1e-8, itmax=200)
    t0 = time.perf_counter()
    xo, ito, erro = po.krylov("CG", b, eps=1e-8, itmax=200)
    _say(f"  CG iters {it} (oracle {ito}, {time.perf_counter() - t0:.1f}s), err {err:.3e} (oracle {erro:.3e})")
    assert err <= 1e-8 and abs(it - ito) <= 1
    xh = xg.cpu().numpy()
    assert true_relres(problems, irp, ja, val, xh, b) <= 1.01e-8
    d64 = rel2(xh, xo)
    del p
    pm = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2, precision="mixed")
    xm, itm, errm = eng.psb_krylov("FCG", None, pm, dev(b), eps=1e-8, itmax=200)
    xmh = xm.cpu().numpy()
    dm = rel2(xmh, xo)
    _say(f"  mixed FCG iters {itm}, err {errm:.3e}; solution vs oracle FP64 CG: FP64 CG {d64:.2e}, mixed {dm:.2e}")
    assert errm <= 1e-8 and abs(itm - ito) <= 1
    assert true_relres(problems, irp, ja, val, xmh, b) <= 1.01e-8
    assert d64 <= 1e-7 and dm <= 1e-7


def test_config2_poisson256_ilu(eng, orc, threads, poisson256):
    """Poisson idim 256, one BJAC/ILU(0) sweep (level-scheduled triangular
    solves on every level): the ILU factors of the fine level bit-identical to
    the oracle's, an application within 1e-12, CG to 1e-8 in the oracle's
    iteration count (+-1) with the true residual <= 1.01e-8."""
    from mld2p4_amd import problems

    irp, ja, val, a = poisson256
    n = len(irp) - 1
    _say(f"[config 2, run B] idim=256 BJAC/ILU(0), oracle threads {threads}")
    p = engine_prec(eng, (irp, ja, val), "BJAC", 1)
    po = oracle_prec(orc, a, "BJAC", 1)
    assert p.nlev == po.nlev
    for what in ("L", "U"):
        _level_arrays_equal(p, po, 0, what)
    assert np.array_equal(p.dv(0), po.dv(0))
    x = np.random.default_rng(257).standard_normal(n)
    r = relmax(p.apply(dev(x)).cpu().numpy(), po.apply(x))
    _say(f"  apply rel diff {r:.2e}")
    assert r <= 1e-12
    b = np.ones(n)
    xg, it, err = eng.psb_krylov("CG", None, p, dev(b), eps=1e-8, itmax=200)
    t0 = time.perf_counter()
    _, ito, _ = po.krylov("CG", b, eps=1e-8, itmax=200)
    _say(f"  CG iters {it} (oracle {ito}, {time.perf_counter() - t0:.1f}s), err {err:.3e}")
    assert err <= 1e-8 and abs(it - ito) <= 1
    assert true_relres(problems, irp, ja, val, xg.cpu().numpy(), b) <= 1.01e-8


def test_config3_diffusion256_mixed(eng, orc, threads):
    """256^3 heterogeneous diffusion (k = exp(2 g), seed 2024), L1-JACOBI 2/2,
    dense FP64 LU of the coarsest level (<= 4096 rows): the FP64 application
    within 1e-12 of the oracle's; FP32-mixed FCG to 1e-8 with the true
    residual <= 1.01e-8 and its solution within 1e-7 (2-norm, relative) of the
    oracle's FP64 CG solution to 1e-12."""
    from mld2p4_amd import problems

    irp, ja, val, bh = problems.diffusion3d(256, seed=2024)
    n = len(irp) - 1
    _say(f"[config 3] diffusion idim=256 n={n}, oracle threads {threads}")
    a = orc.Csr.from_arrays(irp, ja, val)
    po = oracle_prec(orc, a, "L1-JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=4096)
    p = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=4096)
    assert p.nlev == po.nlev
    nc = p.level_info(p.nlev - 1)["n"]
    assert p.level_info(p.nlev - 1)["solver"] == 102 and nc <= 4096
    x = np.random.default_rng(3).standard_normal(n)
    r = relmax(p.apply(dev(x)).cpu().numpy(), po.apply(x))
    _say(f"  FP64 apply rel diff {r:.2e} (coarsest n={nc})")
    assert r <= 1e-12
    t0 = time.perf_counter()
    xo, ito, erro = po.krylov("CG", bh, eps=1e-12, itmax=500)
    _say(f"  oracle FP64 CG to 1e-12: {ito} iterations, err {erro:.2e} ({time.perf_counter() - t0:.1f}s)")
    assert erro <= 1e-12
    x64, it64, err64 = eng.psb_krylov("CG", None, p, dev(bh), eps=1e-8)
    e64 = rel2(x64.cpu().numpy(), xo)
    del p
    pm = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2, coarse="DENSE_LU", csweeps=1, mcs=4096,
                     precision="mixed")
    xm, itm, errm = eng.psb_krylov("FCG", None, pm, dev(bh), eps=1e-8)
    xmh = xm.cpu().numpy()
    em = rel2(xmh, xo)
    _say(f"  FP64 CG {it64} its, error vs oracle {e64:.2e}; mixed FCG {itm} its, error vs oracle {em:.2e}")
    assert errm <= 1e-8
    assert true_relres(problems, irp, ja, val, xmh, bh) <= 1.01e-8
    assert em <= 1e-7


def test_config4a_block192(eng, orc, threads):
    """192^3 Poisson, 32 Gaussian right-hand sides (Philox seed 192), block
    V-cycle: FP64 block application of two columns within 1e-12 of the
    oracle's single-vector applies; FP32 block application within 1e-5; block
    CG with the FP32 preconditioner to 1e-8 with every column's true residual
    <= 1.01e-8."""
    from mld2p4_amd import problems

    irp, ja, val = problems.pde3d_poisson(192)
    n = len(irp) - 1
    _say(f"[config 4a] idim=192 n={n}, 32 RHS")
    a = orc.Csr.from_arrays(irp, ja, val)
    po = oracle_prec(orc, a, "L1-JACOBI", 2)
    rng = np.random.Generator(np.random.Philox(192))
    Bh = rng.standard_normal((32, n))
    B = dev(Bh)
    yo = {j: po.apply(Bh[j]) for j in (0, 31)}
    p64 = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2)
    Y = p64.apply_block(B).cpu().numpy()
    for j in (0, 31):
        r = relmax(Y[j], yo[j])
        _say(f"  FP64 block apply column {j}: rel diff {r:.2e}")
        assert r <= 1e-12
    del p64, Y
    p = engine_prec(eng, (irp, ja, val), "L1-JACOBI", 2, precision="mixed")
    Y = p.apply_block(B).cpu().numpy()
    for j in (0, 31):
        r = relmax(Y[j], yo[j])
        _say(f"  FP32 block apply column {j}: rel diff {r:.2e}")
        assert r <= 1e-5
    del Y
    X, its, errs = p.krylov_block("CG", B, eps=1e-8)
    Xh = X.cpu().numpy()
    import scipy.sparse as sp

    R = Bh - (sp.csr_matrix((val, ja, irp), shape=(n, n)) @ Xh.T).T
    worst = float((np.linalg.norm(R, axis=1) / np.linalg.norm(Bh, axis=1)).max())
    _say(f"  block CG iterations {its.min()}..{its.max()}, worst true residual {worst:.3e}")
    assert errs.max() <= 1e-8 and worst <= 1.01e-8


def test_config4b_group64(eng, orc, threads):
    """64 Poisson systems, idim uniform in [24, 96] (Philox seed 64), one
    N(0,1) right-hand side each: FP64 grouped application of three systems'
    slices within 1e-12 of the oracle's applies, FP32 within 1e-5; grouped
    FCG with FP32 preconditioners to 1e-8 with every system's true residual
    <= 1.01e-8, iterations of all systems in graph launches."""
    from mld2p4_amd import problems

    rng = np.random.Generator(np.random.Philox(64))
    idims = rng.integers(24, 97, size=64)
    _say(f"[config 4b] 64 systems, idim {idims.min()}..{idims.max()}")
    mats = [problems.pde3d_poisson(int(d)) for d in idims]
    systems = [engine_prec(eng, m, "L1-JACOBI", 2, precision="mixed") for m in mats]
    groups = {}
    for prec in ("double", "mixed"):
        g = eng.MldGroup(device=0, nb=1)
        for p in systems:
            g.add(p)
        groups[prec] = g.build(prec)
    ns, nr, off = groups["mixed"].info()
    assert ns == 64 and nr == sum(int(d) ** 3 for d in idims)
    bh = rng.standard_normal(nr)
    b = dev(bh)
    pick = [int(np.argmin(idims)), int(np.argmax(idims)), 17]
    yo = {}
    for s in pick:
        a = orc.Csr.from_arrays(*mats[s])
        yo[s] = oracle_prec(orc, a, "L1-JACOBI", 2).apply(bh[off[s]:off[s + 1]])
    for prec, tol in (("double", 1e-12), ("mixed", 1e-5)):
        y = groups[prec].apply(b).cpu().numpy()
        for s in pick:
            r = relmax(y[off[s]:off[s + 1]], yo[s])
            _say(f"  {prec} group apply, system {s} (idim {idims[s]}): rel diff {r:.2e}")
            assert r <= tol
    x, its, errs, launches = groups["mixed"].krylov("FCG", b, eps=1e-8)
    xh = x.cpu().numpy()
    worst = max(true_relres(problems, *mats[s], xh[off[s]:off[s + 1]], bh[off[s]:off[s + 1]]) for s in range(64))
    _say(f"  group FCG iterations {its.min()}..{its.max()} in {launches} graph launches, worst true residual "
         f"{worst:.3e}")
    assert errs.max() <= 1e-8 and worst <= 1.01e-8


def test_config5_adjoint128(eng, orc, threads):
    """dL/dtheta at idim 128 (theta ~ N(0, 0.5^2) seed 5, b ~ U[-1,1], x* the
    solution at theta* seed 6, L = ||x - x*||^2 / 2):
      * the workload's gradient (FP32-mixed AMG-CG to 1e-10, adjoint solve,
        edge kernel) within 1e-6 (max-norm, relative to max|g|) of the FP64
        reference gradient, the closed form evaluated at the oracle's FP64 CG
        solutions x and lam = A^-T (x - x*) to 1e-13;
      * the FP64 device gradient (solves to 1e-13) within 1e-9 of it;
      * central differences (h = 1e-6) of L on 20 seeded cells, every L from a
        device FP64 solve to 1e-13 with its own hierarchy: |fd - g| <= 1e-5
        max|g|."""
    from mld2p4_amd import autograd as ag
    from mld2p4_amd import problems

    idim = 128
    n = idim ** 3
    th, bh, ths = problems.theta_problem(idim, seed=5)
    _say(f"[config 5] idim={idim}")

    def oracle_solve(theta, rhs):
        irp, ja, val = problems.fv7_matrix(np.exp(theta))
        a = orc.Csr.from_arrays(irp, ja, val)
        po = oracle_prec(orc, a, "L1-JACOBI", 2)
        x, it, err = po.krylov("CG", rhs, eps=1e-13, itmax=500)
        assert err <= 1e-13
        return x, it

    xs, _ = oracle_solve(ths, bh)
    xr, itx = oracle_solve(th, bh)
    lr, itl = oracle_solve(th, xr - xs)  # A(theta) is symmetric: A^T lam = A lam
    from test_adjoint import fv7_grad_ref

    gr = fv7_grad_ref(th, xr, lr)
    gmax = float(np.abs(gr).max())
    _say(f"  oracle solves: {itx} and {itl} iterations; max|g_ref| {gmax:.3e}")
    xsd = dev(xs)
    res = {}
    for prec, eps, tol in (("mixed", 1e-10, 1e-6), ("double", 1e-13, 1e-9)):
        opts = ag.SolveOptions(method="CG", eps=eps, precision=prec)
        t = dev(th).requires_grad_(True)
        x = ag.solve(t, dev(bh), idim, opts)
        loss = 0.5 * torch.sum((x - xsd) ** 2)
        loss.backward()
        g = t.grad.cpu().numpy()
        r = float(np.abs(g - gr).max()) / gmax
        res[prec] = g
        _say(f"  {prec} gradient (solves to {eps:g}: {opts.last}): rel diff vs reference {r:.2e}")
        assert r <= tol
    opts = ag.SolveOptions(method="CG", eps=1e-13, precision="double")

    def L(theta):
        with torch.no_grad():
            xx = ag.solve(dev(theta), dev(bh), idim, opts)
            return float(0.5 * torch.sum((xx - xsd) ** 2))

    cells = np.random.default_rng(2024).choice(n, 20, replace=False)
    h = 1e-6
    worst = 0.0
    g = res["double"]
    for c in cells:
        tp, tm = th.copy(), th.copy()
        tp[c] += h
        tm[c] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        worst = max(worst, abs(fd - g[c]) / gmax)
    _say(f"  central differences on 20 cells: worst |fd - g| / max|g| = {worst:.2e}")
    assert worst <= 1e-5
