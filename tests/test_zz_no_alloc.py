"""mldhip_krylov_noalloc (no device allocation: the Krylov vectors live in the
library's fixed per-GPU workspace) against mldhip_krylov, through the C-ABI:
the same iterations, residual and solution bit for bit, on two calls in a row
(the second on a smaller system, over the workspace contents the first call
left), with the handle's device memory unchanged by the calls; systems above
the workspace's 1048576 rows are refused with -1."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


def build(eng, A, precision):
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                 ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"),
                 ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    p.set_precision(precision)
    return p.build(A)


def noalloc(eng, p, method, b, x, eps=1e-8, itmax=500):
    it = C.c_int(0)
    err = C.c_double(0.0)
    p.use_torch_stream()
    rc = eng.lib().mldhip_krylov_noalloc(p.h, method.encode(), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                         eps, itmax, 2, C.byref(it), C.byref(err))
    p._check(rc, "krylov_noalloc")
    torch.cuda.synchronize()
    return it.value, err.value


@pytest.mark.parametrize("method,precision", [("CG", "double"), ("FCG", "mixed"), ("FCG", "double")])
def test_krylov_noalloc_matches_krylov(eng, orc, method, precision):
    for idim in (96, 40):  # the second call: a smaller system
        _say(f"[krylov noalloc] {method} {precision} idim={idim}")
        a = orc.gen_pde3d(idim)
        A = a.to_scipy()
        n = A.shape[0]
        rng = np.random.default_rng(idim)
        bh = rng.standard_normal(n)
        x0 = rng.standard_normal(n) * 1e-3
        b = torch.from_numpy(bh).to("cuda:0")
        p = build(eng, A, precision)
        before = p.device_bytes()
        x1 = torch.from_numpy(x0.copy()).to("cuda:0")
        it1, err1 = noalloc(eng, p, method, b, x1)
        assert p.device_bytes() == before  # the handle allocated nothing
        x2 = torch.from_numpy(x0.copy()).to("cuda:0")
        _, it2, err2 = eng.psb_krylov(method, A, p, b, x2, eps=1e-8)
        torch.cuda.synchronize()
        assert p.device_bytes() > before  # mldhip_krylov allocates its vectors on first use
        assert it1 == it2 and err1 == err2
        assert torch.equal(x1, x2)
        xh = x1.cpu().numpy()
        assert err1 <= 1e-8
        assert np.linalg.norm(bh - A @ xh) / np.linalg.norm(bh) <= 1.01e-8


def test_krylov_noalloc_refusals(eng, orc):
    _say("[krylov noalloc] refusals")
    A = orc.gen_pde3d(102).to_scipy()  # 1061208 rows > 1048576
    p = build(eng, A, "double")
    b = torch.ones(A.shape[0], dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(b)
    with pytest.raises(eng.MldError) as e:
        noalloc(eng, p, "CG", b, x)
    assert e.value.info == -1 and "1048576" in str(e.value)
    q = build(eng, orc.gen_pde3d(10).to_scipy(), "double")
    b = torch.ones(1000, dtype=torch.float64, device="cuda:0")
    with pytest.raises(eng.MldError) as e:
        noalloc(eng, q, "BICGSTAB", b, torch.zeros_like(b))
    assert e.value.info == -20
