"""Device applications of a hierarchy handed over from outside the engine
(levels and smoother data uploaded through the C-ABI, or loaded from a
prec%dump) and of mld_dmlprec_aply's alpha / beta form, through the C-ABI.

Storage is fixed (SELL: one lane per row, stored order) on both handles so
that the two uploads use the same kernels and the comparison can be bit for
bit; the oracle bar is 1e-12 (max-norm, relative)."""
import numpy as np
import pytest

from test_upload_host import config, export_upload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


SM = {"JACOBI": 1, "BJAC": 2, "L1-JACOBI": 100, "FBGS": 4}


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def relmax(y, yo):
    return float(np.abs(y - yo).max() / max(np.abs(yo).max(), 1e-300))


def oracle_prec(orc, a, smoother, sweeps=2):
    prm = orc.config1_params(smoother=SM[smoother], sweeps_pre=sweeps, sweeps_post=sweeps, min_coarse_size=60)
    return orc.Prec(prm).build(a)


@pytest.mark.parametrize("smoother", ["L1-JACOBI", "BJAC", "FBGS"])
@pytest.mark.parametrize("how", ["levels", "levels+smoothers", "dump", "dump+smoothers"])
def test_handed_over_hierarchy_apply(eng, orc, tmp_path, smoother, how):
    """A hierarchy uploaded level by level (with or without the smoother data)
    or loaded from a dump applies bit-identically to the handle it came from,
    and within 1e-12 of the oracle, 'N' and (Jacobi-type smoothers) 'T'."""
    _say(f"[handed-over hierarchy] {smoother} {how} idim=18")
    a = orc.gen_pde3d(18)
    A = a.to_scipy()
    p = config(eng, smoother, device=0).set_format("sell").build(A)
    q = config(eng, smoother, device=0).set_format("sell")
    if how.startswith("levels"):
        export_upload(p, q, how.endswith("smoothers"))
    else:
        prefix = str(tmp_path / "h")
        p.dump(prefix, istart=1, ac=True, rp=True, smoother=True, solver=True)
        q.load_hierarchy(prefix, A, smoothers=how.endswith("smoothers"))
    if not how.endswith("smoothers"):
        q.smoothers_build()
    po = oracle_prec(orc, a, smoother)
    x = np.random.default_rng(18).standard_normal(A.shape[0])
    for trans in (("N", "T") if smoother != "FBGS" else ("N",)):
        yp = p.apply(dev(x), trans=trans).cpu().numpy()
        yq = q.apply(dev(x), trans=trans).cpu().numpy()
        assert (yp == yq).all(), trans
        assert relmax(yq, po.apply(x, trans=trans)) <= 1e-12, trans
    b = np.ones(A.shape[0])
    _, itp, _ = eng.psb_krylov("CG", A, p, dev(b), eps=1e-8)
    xq, itq, errq = eng.psb_krylov("CG", A, q, dev(b), eps=1e-8)
    assert itp == itq and errq <= 1e-8


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 1e-5)])
def test_mlprec_aply_alpha_beta(eng, orc, precision, tol):
    """y = alpha op(M^-1) x + beta y (mld_dmlprec_aply.f90:327) against the
    oracle's application combined on the host; beta = 0 does not read y
    (a NaN there stays out of the result)."""
    _say(f"[alpha beta] {precision} idim=16")
    a = orc.gen_pde3d(16)
    A = a.to_scipy()
    p = config(eng, "L1-JACOBI", device=0).set_precision(precision).build(A)
    po = oracle_prec(orc, a, "L1-JACOBI")
    rng = np.random.default_rng(16)
    n = A.shape[0]
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    for trans in ("N", "T"):
        mx = po.apply(x, trans=trans)
        for alpha, beta in ((1.0, 0.0), (2.5, 0.0), (1.0, 1.0), (-0.5, 2.0), (0.0, 3.0)):
            y = dev(y0.copy())
            p.mlprec_aply(alpha, dev(x), beta, y, trans=trans)
            ref = alpha * mx + beta * y0
            assert relmax(y.cpu().numpy(), ref) <= tol, (trans, alpha, beta)
        y = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
        p.mlprec_aply(-2.0, dev(x), 0.0, y, trans=trans)
        assert relmax(y.cpu().numpy(), -2.0 * mx) <= tol
