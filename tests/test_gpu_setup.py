"""The hierarchy build's smoothed-aggregation products (P = S Ptilde, A P,
R = P^T, Ac = R A P) on the device against the host products: every level
operator bit-identical (values, columns, row pointers), with the plain and
the filtered smoothed prolongator, and the device products' hierarchy
applied within 1e-12 of the oracle's.  Levels whose rows outgrow the device
accumulator are formed on the host (the coarse levels here)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


def build(eng, A, setup, smoother="L1-JACOBI", theta=0.0, filt=False, mcs=40):
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", smoother), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", theta),
                 ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", mcs), ("COARSE_SOLVE", "BJAC"),
                 ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    if filt:
        p.set("AGGR_FILTER", "FILTER")
    p.set_setup(setup)
    return p.build(A)


def same(m1, m2):
    return (m1.shape == m2.shape and np.array_equal(m1.indptr, m2.indptr) and np.array_equal(m1.indices, m2.indices)
            and np.array_equal(m1.data, m2.data))


def assert_same_hierarchy(pd, ph):
    assert pd.nlev == ph.nlev
    for l in range(pd.nlev):
        for w in (["A"] if l == 0 else ["A", "P", "R"]):
            assert same(pd.level_csr(l, w), ph.level_csr(l, w)), (l, w)
        if l > 0:
            assert pd.omega(l) == ph.omega(l)
            assert np.array_equal(pd.aggr(l), ph.aggr(l))


@pytest.mark.parametrize("idim", [12, 33, 64])
def test_device_products_poisson(eng, orc, idim):
    _say(f"[device setup] Poisson idim={idim}")
    a = orc.gen_pde3d(idim)
    A = a.to_scipy()
    pd = build(eng, A, "device")
    ph = build(eng, A, "host")
    assert_same_hierarchy(pd, ph)
    po = orc.Prec(orc.config1_params(smoother=100, sweeps_pre=2, sweeps_post=2, min_coarse_size=40)).build(a)
    x = np.random.default_rng(idim).standard_normal(A.shape[0])
    y = pd.apply(torch.from_numpy(x).to("cuda:0")).cpu().numpy()
    yo = po.apply(x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()


@pytest.mark.parametrize("filt", [False, True])
def test_device_products_diffusion(eng, orc, filt):
    """Heterogeneous diffusion with a strength threshold (and the filtered
    prolongator): the device products equal the host's."""
    _say(f"[device setup] diffusion idim=20 filter={filt}")
    from mld2p4_amd import problems

    irp, ja, val, _ = problems.diffusion3d(20, seed=11, passes=0)
    pd = build(eng, (irp, ja, val), "device", theta=0.1, filt=filt, mcs=20)
    ph = build(eng, (irp, ja, val), "host", theta=0.1, filt=filt, mcs=20)
    assert pd.nlev >= 3
    assert_same_hierarchy(pd, ph)
