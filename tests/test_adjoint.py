"""Differentiation of the AMG solve: adjoint Krylov (hierarchy applied with
trans = 'T'), the device edge-product gradient of the 7-point finite-volume
operator, and the torch.autograd harness.

The host gradient formula below (fv7_grad_ref) is the checker.  It is pinned
on the CPU against central finite differences of the bilinear form
lam^T A(theta) x, assembled by problems.fv7_matrix (an independent code path:
harmonic face means, boundary faces), and the whole chain dL/dtheta against
central finite differences of L(theta) with scipy's direct solver.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def fv7_grad_ref(theta, x, lam):
    """-lam^T (dA/dtheta_c) x for every cell c (numpy, vectorised)."""
    n = theta.size
    idim = int(round(n ** (1.0 / 3.0)))
    k = np.exp(theta).reshape(idim, idim, idim)
    X = x.reshape(idim, idim, idim)
    L = lam.reshape(idim, idim, idim)
    g = np.zeros_like(k)
    for ax in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        kl, kh = k[lo], k[hi]
        prod = (L[lo] - L[hi]) * (X[lo] - X[hi])
        den2 = (kl + kh) ** 2
        g[lo] += 2.0 * kl * kh * kh / den2 * prod
        g[hi] += 2.0 * kh * kl * kl / den2 * prod
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[ax], last[ax] = 0, -1
        first, last = tuple(first), tuple(last)
        g[first] += k[first] * L[first] * X[first]
        g[last] += k[last] * L[last] * X[last]
    return -(g * idim * idim).reshape(n)


def fv7_sp(theta):
    from mld2p4_amd import problems

    irp, ja, val = problems.fv7_matrix(np.exp(theta))
    n = theta.size
    return sp.csr_matrix((val, ja, irp), shape=(n, n))


def test_grad_formula_vs_bilinear_fd():
    """Central FD (h = 1e-5) of q(theta) = lam^T A(theta) x on interior,
    face, edge and corner cells matches the closed form."""
    idim = 5
    rng = np.random.default_rng(1)
    th = 0.7 * rng.standard_normal(idim ** 3)
    x = rng.standard_normal(idim ** 3)
    lam = rng.standard_normal(idim ** 3)
    g = fv7_grad_ref(th, x, lam)
    h = 1e-5
    cells = [0, 2, 12, 31, 62, idim ** 3 - 1, (2 * idim + 2) * idim + 2]
    for c in cells:
        tp, tm = th.copy(), th.copy()
        tp[c] += h
        tm[c] -= h
        fd = (lam @ (fv7_sp(tp) @ x) - lam @ (fv7_sp(tm) @ x)) / (2 * h)
        assert abs(fd + g[c]) <= 1e-7 * max(1.0, abs(fd)), (c, fd, -g[c])


def test_grad_chain_vs_fd_scipy():
    """dL/dtheta = -lam^T dA/dtheta x with A^T lam = x - x* (L = ||x - x*||^2/2):
    checked against central FD of L with direct solves (h = 1e-6, 20 cells)."""
    from mld2p4_amd import problems

    idim = 6
    th, b, ths = problems.theta_problem(idim, seed=3)
    xs = spla.spsolve(fv7_sp(ths).tocsc(), b)

    def loss(t):
        x = spla.spsolve(fv7_sp(t).tocsc(), b)
        return 0.5 * np.sum((x - xs) ** 2), x

    L0, x = loss(th)
    lam = spla.spsolve(fv7_sp(th).T.tocsc(), x - xs)
    g = fv7_grad_ref(th, x, lam)
    cells = np.random.default_rng(7).choice(idim ** 3, 20, replace=False)
    h = 1e-6
    for c in cells:
        tp, tm = th.copy(), th.copy()
        tp[c] += h
        tm[c] -= h
        fd = (loss(tp)[0] - loss(tm)[0]) / (2 * h)
        assert abs(fd - g[c]) <= 1e-6 * np.abs(g).max(), (c, fd, g[c])


def test_adjoint_symbols_exported(eng):
    lib = eng.lib()
    assert hasattr(lib, "mldhip_krylov_adjoint") and hasattr(lib, "mldhip_fv7_grad_theta")


# ----------------------------------------------------------------------- GPU
def _say(msg):
    print(msg, flush=True)


def dev(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def convection_matrix(idim, beta=40.0):
    """Nonsymmetric test operator: fv7 diffusion plus first-order upwind
    convection along x (an M-matrix)."""
    from mld2p4_amd import problems

    th, _, _ = problems.theta_problem(idim, seed=9)
    A = fv7_sp(th)
    n = idim ** 3
    g = np.arange(n)
    has_xm = g >= idim * idim
    rows = np.concatenate([g, g[has_xm]])
    cols = np.concatenate([g, g[has_xm] - idim * idim])
    vals = np.concatenate([np.full(n, beta * idim), np.full(has_xm.sum(), -beta * idim)])
    A = (A + sp.csr_matrix((vals, (rows, cols)), shape=(n, n))).tocsr()
    A.sort_indices()
    return A


@pytest.mark.gpu
@pytest.mark.parametrize("idim", [13, 20])
def test_fv7_grad_kernel(eng, idim):
    _say(f"[fv7 grad kernel] idim={idim}")
    import torch

    rng = np.random.default_rng(idim)
    th, x, lam = (rng.standard_normal(idim ** 3) for _ in range(3))
    gr = fv7_grad_ref(th, x, lam)
    for work in (True, None):  # exp(theta) precomputed / formed in place
        g = eng.fv7_grad_theta(idim, dev(th), dev(x), dev(lam), work=work)
        torch.cuda.synchronize()
        assert np.abs(g.cpu().numpy() - gr).max() <= 1e-13 * np.abs(gr).max()


@pytest.mark.gpu
def test_apply_transpose_nonsymmetric(eng, orc):
    """'T' apply of a hierarchy built on a nonsymmetric matrix vs the oracle."""
    _say("[apply T nonsym] idim=12")
    A = convection_matrix(12)
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                 ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"), ("COARSE_SUBSOLVE", "ILU"),
                 ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    p.build(A)
    po = orc.Prec(orc.config1_params(sweeps_pre=2, sweeps_post=2)).build(orc.Csr.from_scipy(A))
    assert p.nlev == po.nlev >= 2
    x = np.random.default_rng(3).standard_normal(A.shape[0])
    for trans in ("N", "T"):
        y = p.apply(dev(x), trans=trans).cpu().numpy()
        yo = po.apply(x, trans=trans)
        assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max(), trans
    # and the two differ: the test exercises the transposed operators
    assert np.abs(p.apply(dev(x), trans="T").cpu().numpy() - p.apply(dev(x)).cpu().numpy()).max() > 1e-6


@pytest.mark.gpu
def test_krylov_adjoint_nonsymmetric(eng):
    """BiCGStab on A^T lam = g with the hierarchy applied transposed, for a
    nonsymmetric A (CG-type methods need a symmetric A)."""
    _say("[krylov adjoint] BICGSTAB idim=14")
    A = convection_matrix(14)
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                 ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"), ("COARSE_SUBSOLVE", "ILU"),
                 ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    p.build(A)
    g = np.random.default_rng(5).standard_normal(A.shape[0])
    lam, it, err = eng.psb_krylov_adjoint("BICGSTAB", p, dev(g), eps=1e-11, itmax=2000)
    assert err <= 1e-11
    lr = lam.cpu().numpy()
    assert np.linalg.norm(A.T @ lr - g) <= 1.01e-11 * np.linalg.norm(g)
    # forward solve with the same handle still works afterwards
    xf, itf, errf = eng.psb_krylov("BICGSTAB", A, p, dev(g), eps=1e-11, itmax=2000)
    assert np.linalg.norm(A @ xf.cpu().numpy() - g) <= 1.01e-11 * np.linalg.norm(g)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["CG", "FCG"])
def test_krylov_adjoint_symmetric(eng, method):
    """Symmetric diffusion operator: the adjoint solve (transposed hierarchy)
    solves A^T lam = g = A lam to the tolerance."""
    _say(f"[krylov adjoint] {method} diffusion idim=16")
    from mld2p4_amd import problems

    th, b, _ = problems.theta_problem(16, seed=4)
    A = fv7_sp(th)
    p = eng.MldPrec(device=0).init("ML")
    for k, v in [("SMOOTHER_TYPE", "L1-JACOBI"), ("SMOOTHER_SWEEPS", 2), ("AGGR_THRESH", 0.0),
                 ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"), ("COARSE_SUBSOLVE", "ILU"),
                 ("COARSE_SWEEPS", 4)]:
        p.set(k, v)
    p.build(A)
    lam, it, err = eng.psb_krylov_adjoint(method, p, dev(b), eps=1e-10)
    assert err <= 1e-10
    assert np.linalg.norm(A.T @ lam.cpu().numpy() - b) <= 1.01e-10 * np.linalg.norm(b)
    _, itf, _ = eng.psb_krylov(method, A, p, dev(b), eps=1e-10)
    assert abs(it - itf) <= 1


def _grad_setup(idim):
    from mld2p4_amd import problems

    th, b, ths = problems.theta_problem(idim, seed=5)
    xs = spla.spsolve(fv7_sp(ths).tocsc(), b)
    return th, b, xs


@pytest.mark.gpu
def test_autograd_vs_fp64_reference_gradient(eng):
    """Autograd gradient (device forward + adjoint solve + edge kernel) vs
    the FP64 reference gradient (direct solves + closed form):
    max-norm relative difference <= 1e-9 with solves to 1e-13."""
    _say("[autograd vs reference] idim=16")
    import torch
    from mld2p4_amd import autograd as ag

    idim = 16
    th, b, xs = _grad_setup(idim)
    opts = ag.SolveOptions(method="CG", eps=1e-13)
    t = dev(th).requires_grad_(True)
    bt = dev(b).requires_grad_(True)
    x = ag.solve(t, bt, idim, opts)
    loss = 0.5 * torch.sum((x - dev(xs)) ** 2)
    loss.backward()
    A = fv7_sp(th)
    xr = spla.spsolve(A.tocsc(), b)
    lr = spla.spsolve(A.T.tocsc(), xr - xs)
    gr = fv7_grad_ref(th, xr, lr)
    gt = t.grad.cpu().numpy()
    print(f"iters fwd {opts.last['forward_iters']} adj {opts.last['adjoint_iters']}; "
          f"rel diff {np.abs(gt - gr).max() / np.abs(gr).max():.2e}")
    assert np.abs(gt - gr).max() <= 1e-9 * np.abs(gr).max()
    assert np.abs(bt.grad.cpu().numpy() - lr).max() <= 1e-9 * np.abs(lr).max()


@pytest.mark.gpu
def test_autograd_vs_central_fd(eng):
    """Central finite differences (h = 1e-6) of L(theta) = ||x - x*||^2/2 on
    20 seeded cells, every forward solve on the device (solves to 1e-13):
    |g_fd - g| <= 1e-5 max|g|."""
    _say("[autograd vs FD] idim=12")
    import torch
    from mld2p4_amd import autograd as ag

    idim = 12
    th, b, xs = _grad_setup(idim)
    opts = ag.SolveOptions(method="CG", eps=1e-13)
    t = dev(th).requires_grad_(True)
    x = ag.solve(t, dev(b), idim, opts)
    loss = 0.5 * torch.sum((x - dev(xs)) ** 2)
    loss.backward()
    g = t.grad.cpu().numpy()
    xsd = dev(xs)

    def L(th_):
        with torch.no_grad():
            xx = ag.solve(dev(th_), dev(b), idim, opts)
            return float(0.5 * torch.sum((xx - xsd) ** 2))

    cells = np.random.default_rng(2024).choice(idim ** 3, 20, replace=False)
    h = 1e-6
    worst = 0.0
    for c in cells:
        tp, tm = th.copy(), th.copy()
        tp[c] += h
        tm[c] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        worst = max(worst, abs(fd - g[c]) / np.abs(g).max())
    print(f"worst FD difference / max|g| = {worst:.2e}")
    assert worst <= 1e-5


@pytest.mark.gpu
def test_autograd_mixed_vs_fp64_reference(eng):
    """The differentiation workload's solver settings (FP32-mixed AMG-CG to
    1e-10, forward and adjoint): gradient within 1e-6 (max-norm, relative)
    of the FP64 reference gradient."""
    _say("[autograd mixed vs reference] idim=20")
    import torch
    from mld2p4_amd import autograd as ag

    idim = 20
    th, b, xs = _grad_setup(idim)
    opts = ag.SolveOptions(method="CG", eps=1e-10, precision="mixed")
    t = dev(th).requires_grad_(True)
    x = ag.solve(t, dev(b), idim, opts)
    loss = 0.5 * torch.sum((x - dev(xs)) ** 2)
    loss.backward()
    A = fv7_sp(th)
    xr = spla.spsolve(A.tocsc(), b)
    lr = spla.spsolve(A.T.tocsc(), xr - xs)
    gr = fv7_grad_ref(th, xr, lr)
    rel = np.abs(t.grad.cpu().numpy() - gr).max() / np.abs(gr).max()
    print(f"iters fwd {opts.last['forward_iters']} adj {opts.last['adjoint_iters']}; rel diff {rel:.2e}")
    assert rel <= 1e-6
