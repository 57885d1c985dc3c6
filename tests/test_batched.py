"""Batched forms on the GPU, through the C-ABI:

* block V-cycle (32 columns, sparse-times-dense products): every column of
  the block apply equals the oracle's single-vector apply (FP64: <= 1e-12
  relative; FP32 preconditioner: <= 1e-5), and block CG converges per column
  like the single-vector CG;
* grouped systems (different sizes and level counts, one union V-cycle, one
  hipGraph for the Krylov iterations): each system's slice of the group
  apply equals the oracle's apply of that system (<= 1e-12), and each
  system's CG converges to the requested tolerance with the single-system
  iteration count (+-1).

The coarsest solve of the batched engines is the explicit operator of the
coarse smoother (same operator as the oracle's substitution; rounding
differs, within the tolerances above).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def relmax(y, yo):
    return float(np.abs(y - yo).max() / max(np.abs(yo).max(), 1e-300))


SETTINGS = [("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0), ("AGGR_PROL", "SMOOTHED"),
            ("COARSE_SOLVE", "BJAC"), ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]


def eng_prec(eng, A, mcs=200, precision="double", extra=()):
    p = eng.MldPrec(device=0).init("ML")
    for k, v in list(SETTINGS) + [("MIN_COARSE_SIZE", mcs)] + list(extra):
        p.set(k, v)
    p.set_precision(precision)
    return p.build(A)


def orc_prec(orc, a, mcs=200, smoother=None, sweeps=2):
    prm = orc.config1_params(smoother=orc.SM_L1JAC if smoother is None else smoother, sweeps_pre=sweeps,
                             sweeps_post=sweeps, min_coarse_size=mcs)
    return orc.Prec(prm).build(a)


@pytest.mark.parametrize("nrhs", [32, 5])
def test_block_apply_fp64(eng, orc, nrhs):
    _say(f"[block apply fp64] idim=12 nrhs={nrhs}")
    a = orc.gen_pde3d(12)
    A = a.to_scipy()
    p = eng_prec(eng, A)
    po = orc_prec(orc, a)
    n = A.shape[0]
    X = np.random.default_rng(nrhs).standard_normal((nrhs, n))
    Y = p.apply_block(dev(X)).cpu().numpy()
    for j in range(nrhs):
        assert relmax(Y[j], po.apply(X[j])) <= 1e-12, j


def test_block_apply_jacobi_w_cycle(eng, orc):
    """Point Jacobi, 3 sweeps (the multi-sweep path), W-cycle."""
    _say("[block apply jacobi W] idim=10")
    a = orc.gen_pde3d(10)
    A = a.to_scipy()
    p = eng_prec(eng, A, mcs=40, extra=[("SMOOTHER_TYPE", "JACOBI"), ("SMOOTHER_SWEEPS", 3), ("ML_CYCLE", "WCYCLE")])
    prm = orc.config1_params(smoother=orc.SM_JAC, sweeps_pre=3, sweeps_post=3, min_coarse_size=40, ml_cycle=4)
    po = orc.Prec(prm).build(a)
    assert p.nlev == po.nlev >= 3
    X = np.random.default_rng(1).standard_normal((32, A.shape[0]))
    Y = p.apply_block(dev(X)).cpu().numpy()
    for j in (0, 7, 31):
        assert relmax(Y[j], po.apply(X[j])) <= 1e-12, j


def test_block_apply_mixed(eng, orc):
    _say("[block apply mixed] idim=16")
    a = orc.gen_pde3d(16)
    A = a.to_scipy()
    p = eng_prec(eng, A, precision="mixed")
    po = orc_prec(orc, a)
    X = np.random.default_rng(3).standard_normal((32, A.shape[0]))
    Y = p.apply_block(dev(X)).cpu().numpy()
    for j in range(0, 32, 5):
        assert relmax(Y[j], po.apply(X[j])) <= 1e-5, j


@pytest.mark.parametrize("split", ["1,0", "1,1", "2,1", "4,0", "4,1", "4,2", "2,2"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 1e-5)])
def test_block_apply_split(eng, orc, split, precision, tol, monkeypatch):
    """Every columns-per-lane split and row-block order of the block products
    (DIA fine level, CSR coarse levels; at idim=20 with one column per lane
    the DIA chunk is 6 blocks, with a partial last period)."""
    _say(f"[block apply split] {split} {precision} idim=20")
    monkeypatch.setenv("MLD_BLOCK_SPLIT", split)
    a = orc.gen_pde3d(20)
    A = a.to_scipy()
    p = eng_prec(eng, A, precision=precision)
    po = orc_prec(orc, a)
    X = np.random.default_rng(11).standard_normal((32, A.shape[0]))
    Y = p.apply_block(dev(X)).cpu().numpy()
    for j in (0, 9, 22, 31):
        assert relmax(Y[j], po.apply(X[j])) <= tol, j


def test_block_split_rejects_bad_value(eng, orc, monkeypatch):
    monkeypatch.setenv("MLD_BLOCK_SPLIT", "3,0")
    a = orc.gen_pde3d(8)
    p = eng_prec(eng, a.to_scipy())
    with pytest.raises(Exception, match="MLD_BLOCK_SPLIT"):
        p.apply_block(dev(np.ones((32, 512))))


@pytest.mark.parametrize("method,precision", [("CG", "double"), ("FCG", "mixed")])
def test_block_krylov(eng, orc, method, precision):
    _say(f"[block krylov] {method} {precision} idim=20")
    a = orc.gen_pde3d(20)
    A = a.to_scipy()
    p = eng_prec(eng, A, precision=precision)
    n = A.shape[0]
    B = np.random.default_rng(7).standard_normal((32, n))
    X, it, er = p.krylov_block(method, dev(B), eps=1e-8)
    Xh = X.cpu().numpy()
    assert (er <= 1e-8).all()
    for j in range(32):
        assert np.linalg.norm(B[j] - A @ Xh[j]) <= 1.01e-8 * np.linalg.norm(B[j])
    # per-column iteration counts as the single-vector solver
    for j in (0, 13):
        _, it1, _ = eng.psb_krylov(method, A, p, dev(B[j]), eps=1e-8)
        assert abs(int(it[j]) - it1) <= 1, (j, it[j], it1)


IDIMS = [9, 14, 6, 12, 17, 7]


def build_group(eng, orc, precision="double", mcs=60):
    systems, oracles, mats = [], [], []
    for idim in IDIMS:
        a = orc.gen_pde3d(idim)
        A = a.to_scipy()
        systems.append(eng_prec(eng, A, mcs=mcs, precision=precision))
        oracles.append(orc_prec(orc, a, mcs=mcs))
        mats.append(A)
    g = eng.MldGroup(device=0, nb=1)
    for p in systems:
        g.add(p)
    g.build(precision)
    return g, systems, oracles, mats


def test_group_apply_fp64(eng, orc):
    _say(f"[group apply fp64] idims={IDIMS}")
    g, systems, oracles, mats = build_group(eng, orc)
    nlevs = [p.nlev for p in systems]
    assert len(set(nlevs)) >= 2, nlevs  # uneven hierarchies
    ns, nr, off = g.info()
    assert ns == len(IDIMS) and nr == sum(i ** 3 for i in IDIMS)
    x = np.random.default_rng(11).standard_normal(nr)
    y = g.apply(dev(x)).cpu().numpy()
    for s, po in enumerate(oracles):
        sl = slice(off[s], off[s + 1])
        assert relmax(y[sl], po.apply(x[sl])) <= 1e-12, s


@pytest.mark.parametrize("method,precision,chunk", [("CG", "double", None), ("FCG", "mixed", "1")])
def test_group_krylov(eng, orc, method, precision, chunk, monkeypatch):
    _say(f"[group krylov] {method} {precision} chunk={chunk}")
    if chunk:
        monkeypatch.setenv("MLD_GRAPH_CHUNK", chunk)
    g, systems, oracles, mats = build_group(eng, orc, precision)
    ns, nr, off = g.info()
    b = np.random.default_rng(5).standard_normal(nr)
    x, it, er, launches = g.krylov(method, dev(b), eps=1e-8)
    xh = x.cpu().numpy()
    assert (er <= 1e-8).all(), er
    assert launches >= 1
    for s, A in enumerate(mats):
        sl = slice(off[s], off[s + 1])
        assert np.linalg.norm(b[sl] - A @ xh[sl]) <= 1.01e-8 * np.linalg.norm(b[sl]), s
        _, it1, _ = eng.psb_krylov(method, A, systems[s], dev(b[sl]), eps=1e-8)
        assert abs(int(it[0, s]) - it1) <= 1, (s, it[0, s], it1)
    # a second solve runs as one graph launch of as many iterations as the
    # first one took, with the same result
    x2, it2, er2, l2 = g.krylov(method, dev(b), eps=1e-8)
    assert (x2.cpu().numpy() == xh).all() and (it2 == it).all()
    assert l2 == 1


def test_group_block_columns(eng, orc):
    """A group with 32 columns per system (grouped block V-cycle)."""
    _say("[group nb=32 apply]")
    systems, oracles = [], []
    for idim in (8, 11):
        a = orc.gen_pde3d(idim)
        systems.append(eng_prec(eng, a.to_scipy(), mcs=40))
        oracles.append(orc_prec(orc, a, mcs=40))
    g = eng.MldGroup(device=0, nb=32)
    for p in systems:
        g.add(p)
    g.build("double")
    ns, nr, off = g.info()
    X = np.random.default_rng(2).standard_normal((32, nr))
    Y = g.apply(dev(X)).cpu().numpy()
    for s, po in enumerate(oracles):
        sl = slice(off[s], off[s + 1])
        for j in (0, 31):
            assert relmax(Y[j, sl], po.apply(X[j, sl])) <= 1e-12, (s, j)
