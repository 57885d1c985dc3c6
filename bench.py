#!/usr/bin/env python3
"""Benchmark: AMG-preconditioned CG time to relative residual 1e-8 on one MI355X.

Workload (the FP64 GPU case): 3-D Poisson, 7-point stencil of the
reference's pdegen driver at idim=256 (16 777 216 unknowns, 117 047 296
nonzeros), FP64 values, int32 indices, b = 1, x0 = 0.  Preconditioner:
decoupled smoothed aggregation (threshold 0, omega = 4/(3 ||D^-1 A||_inf)),
V-cycle with 2 pre / 2 post l1-Jacobi sweeps, MIN_COARSE_SIZE 200,
coarsest level BJAC/ILU(0) (4 sweeps = one ILU solve on one device).
Outer CG (istop 2) to ||r||/||b|| <= 1e-8.

A step is one complete solve from x0 = 0 (the hierarchy is built and
uploaded once before timing; inputs resident in HBM).  value = mean seconds
per solve over the timed steps.

Reported beside it, in the same JSON line:
  mixed      : FP32 hierarchy inside FP64 FCG, same problem and hierarchy;
  run_b      : block-Jacobi ILU(0), 1 sweep, level-scheduled solves;
  diffusion  : 256^3 heterogeneous diffusion, FP32-mixed FCG with a dense
               FP64 LU (MFMA) coarsest level vs the FP64 path, error vs a
               FP64 reference solution;
  batched    : 32 right-hand sides at idim=192 through the block V-cycle,
               and 64 independent systems (idim uniform in [24, 96]) in one
               graph-captured solve, each against one-at-a-time solves;
  adjoint    : dL/dtheta at idim=128 (forward + adjoint solve + the
               edge-product gradient kernel);
  cpu_baseline: the oracle (sequential CPU restatement) on a bounded sample.
"""
from __future__ import annotations

import argparse
import gc
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def _engine():
    name = "mld2p4_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "mld2p4-2_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


BASE = [("AGGR_THRESH", 0.0), ("AGGR_PROL", "SMOOTHED")]


def make_prec(eng, smoother="L1-JACOBI", sweeps=2, precision="double", mcs=200, coarse="BJAC"):
    p = eng.MldPrec(device=0).init("ML")
    p.set("SMOOTHER_TYPE", smoother)
    p.set("SMOOTHER_SWEEPS", sweeps)
    for k, v in BASE:
        p.set(k, v)
    p.set("MIN_COARSE_SIZE", mcs)
    p.set("COARSE_SOLVE", coarse)
    if coarse == "BJAC":
        p.set("COARSE_SUBSOLVE", "ILU")
        p.set("COARSE_SWEEPS", 4)
    else:
        p.set("COARSE_SWEEPS", 1)
    p.set_precision(precision)
    return p


def level_table(p):
    levels = []
    for l in range(p.nlev):
        i = p.level_info(l)
        d = dict(n=i["n"], nnz_a=i["nnz_a"], nnz_p=i["nnz_p"], nnz_r=i["nnz_r"], nnz_nd=i["nnz_nd"],
                 solver=i["solver"], sweeps_pre=i["sweeps_pre"], sweeps_post=i["sweeps_post"])
        fa = p.op_format(l, "A")
        d["csr_a"] = fa["fmt"] == "CSR"
        d["fmt_a"] = fa["fmt"]
        if l > 0:
            d["fmt_p"] = p.op_format(l, "P")["fmt"]
            d["fmt_r"] = p.op_format(l, "R")["fmt"]
        if i["nnz_nd"] > 0:
            d["csr_nd"] = p.op_format(l, "ND")["fmt"] == "CSR"
            d["fmt_nd"] = p.op_format(l, "ND")["fmt"]
        if i["solver"] == 7:
            d["nnz_lu"] = p.level_csr(l, "L").nnz + p.level_csr(l, "U").nnz
        levels.append(d)
    return levels


def solve_timed(eng, torch, p, method, b, eps, steps, warmup, itmax=1000):
    x = torch.zeros_like(b)
    for _ in range(warmup):
        x.zero_()
        eng.psb_krylov(method, None, p, b, x, eps=eps, itmax=itmax)
    torch.cuda.synchronize()
    times, its, errs = [], [], []
    for _ in range(steps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, it, err = eng.psb_krylov(method, None, p, b, x, eps=eps, itmax=itmax)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        its.append(it)
        errs.append(err)
    return x, times, its, errs


def copy_bandwidth(torch, reps=10):
    """Device-to-device copy of 2 GiB (torch's copy_, the HIP runtime's copy
    kernel), read + write bytes per second, reported beside the roofline as the
    rate of a plain streaming copy on this GPU."""
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * n * 8 / t / 1e9


def time_vcycle(torch, p, n, reps=20):
    xr = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(xr)
    p.apply(xr, y)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        p.apply(xr, y)
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def release(torch):
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def cpu_info():
    """Host CPU model, nproc, the CPUs this process may use and the thread
    count of an all-cores run (OMP_NUM_THREADS when set, else every usable
    CPU)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count()
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(omp) if omp.isdigit() and int(omp) > 0 else avail
    return {"cpu_model": model, "nproc": nproc, "cpus_available": avail, "threads_all_cores": min(threads, avail)}


def _oracle():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O

    return O


def cpu_baseline(idim, gpu_iters, sample_iters=5):
    """The benchmark workload on the oracle (the CPU restatement; its results
    do not depend on the thread count) on this host: hierarchy + smoothers
    build once, then CG iterations on one core and on all cores, one warm-up
    iteration and the median of `sample_iters` timed ones (steady_clock),
    scaled to the GPU's iteration count."""
    O = _oracle()
    info = cpu_info()
    a = O.gen_pde3d(idim)
    prm = O.config1_params(smoother=O.SM_L1JAC, sweeps_pre=2, sweeps_post=2)
    # the thread count is process-wide OpenMP state (the engine's host build
    # shares it): restored on the way out
    threads0 = O.set_threads(0)
    O.set_threads(info["threads_all_cores"])
    po = O.Prec(prm).build(a)
    tb = po.build_s
    b = np.ones(a.shape[0])
    per_it = {}
    for cores in (1, info["threads_all_cores"]):
        O.set_threads(cores)
        _, it, _, its, _ = po.krylov("CG", b, eps=1e-30, itmax=1 + sample_iters, timings=True)
        per_it[cores] = float(np.median(its[1:]))
    O.set_threads(threads0)
    one, allc = per_it[1], per_it[info["threads_all_cores"]]
    return {
        "value": round(one * gpu_iters, 3),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (C++ restatement) at idim={idim}: hierarchy + smoothers build {tb:.1f}s (not in value); "
                   f"CG iterations: 1 warm-up + median of {sample_iters} timed (steady_clock) = {one:.3f}s/iter on "
                   f"1 core, x {gpu_iters} GPU iterations"),
        "all_cores": {"value": round(allc * gpu_iters, 3), "cores": info["threads_all_cores"],
                      "s_per_iter": round(allc, 4),
                      "note": "row loops of products and vector updates over OpenMP threads; dots and ILU "
                              "substitutions sequential"},
        "setup_s": round(tb, 2),
        **info,
    }


def section_c1(eng, torch, problems):
    """The reference's CPU-runnable case (idim 40, JACOBI 1/1, coarsest 4
    sweeps of BJAC/ILU(0), CG to 1e-6): the oracle on one core against the
    GPU path, setup (hierarchy_build + smoothers_build) and solve each as the
    median of 5 after one warm-up."""
    O = _oracle()
    threads0 = O.set_threads(0)  # process-wide OpenMP state, restored below
    O.set_threads(1)
    idim = 40
    a = O.gen_pde3d(idim)
    b = np.ones(a.shape[0])
    setups, solves, ito = [], [], 0
    for k in range(6):
        po = O.Prec(O.config1_params()).build(a)
        _, ito, erro, _, tot = po.krylov("CG", b, eps=1e-6, timings=True)
        if k:
            setups.append(po.build_s)
            solves.append(tot)
    O.set_threads(threads0)
    irp, ja, val = problems.pde3d_poisson(idim)
    bd = torch.ones(len(irp) - 1, dtype=torch.float64, device="cuda:0")
    gsetups, gsolves, itg = [], [], 0
    for k in range(6):
        p = eng.MldPrec(device=0).init("ML")
        for key, v in [("SMOOTHER_TYPE", "JACOBI"), ("SMOOTHER_SWEEPS", 1), ("AGGR_THRESH", 0.0),
                       ("AGGR_PROL", "SMOOTHED"), ("MIN_COARSE_SIZE", 200), ("COARSE_SOLVE", "BJAC"),
                       ("COARSE_SUBSOLVE", "ILU"), ("COARSE_SWEEPS", 4)]:
            p.set(key, v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.build((irp, ja, val))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x = torch.zeros_like(bd)
        _, itg, errg = eng.psb_krylov("CG", None, p, bd, x, eps=1e-6)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if k:
            gsetups.append(t1 - t0)
            gsolves.append(t2 - t1)
        del p
    cpu_solve, gpu_solve = float(np.median(solves)), float(np.median(gsolves))
    return {"workload": "3-D Poisson idim=40 (n=64000), b=1, x0=0; JACOBI 1/1, theta 0, smoothed prolongator, "
                        "MIN_COARSE_SIZE 200, coarsest 4 sweeps of BJAC/ILU(0); CG to 1e-6",
            "cpu": {"setup_s": round(float(np.median(setups)), 4), "solve_s": round(cpu_solve, 4), "iters": ito,
                    "relres": erro, "cores": 1, "kind": "port"},
            "gpu": {"setup_s": round(float(np.median(gsetups)), 4), "solve_s": round(gpu_solve, 5), "iters": itg,
                    "relres": errg},
            "solve_speedup_gpu_vs_cpu": round(cpu_solve / gpu_solve, 2),
            "timing": "1 warm-up + median of 5 (CPU: steady_clock inside the oracle; GPU: host clock around "
                      "synchronized calls)"}


# ----------------------------------------------------------------- sections
def section_run_b(eng, torch, problems, irp, ja, val, n, steps):
    """Block-Jacobi ILU(0), one sweep, on the same Poisson problem."""
    p = make_prec(eng, "BJAC", 1)
    p.set("SUB_SOLVE", "ILU")
    t0 = time.perf_counter()
    p.build((irp, ja, val))
    t_b = time.perf_counter() - t0
    b = torch.ones(n, dtype=torch.float64, device="cuda:0")
    _, tm, it, er = solve_timed(eng, torch, p, "CG", b, 1e-8, steps, 1)
    t_vc = time_vcycle(torch, p, n, reps=5)
    # dependency depth: level sets of each ILU(0) solve per level, times the
    # solves per V-cycle (pre + post sweeps, each an L and a U solve)
    levels, steps_vc, steps_sched = [], 0, 0
    for l in range(p.nlev):
        i = p.level_info(l)
        dl, du = p.tri_levels(l)
        apps = i["sweeps_pre"] + (i["sweeps_post"] if l < p.nlev - 1 else 0)
        levels.append({"n": i["n"], "levels_lower": dl, "levels_upper": du, "ilu_applications": apps})
        steps_vc += apps * (dl + du)
        if l > 0:
            steps_sched += apps * (dl + du)
    # the fine level's factors are those of the 7-point grid: grid
    # substitutions (one wavefront per 8 x 8 tile of grid lines) instead of its
    # level sets; one application timed both ways
    fine_grid_ms = p.time_step(0, "pre", 5)
    os.environ["MLD_TRSV_GRID"] = "0"
    try:
        fine_sched_ms = p.time_step(0, "pre", 5)
    finally:
        os.environ.pop("MLD_TRSV_GRID", None)
    del p
    release(torch)
    floor_us = trsv_chain_step(eng, torch)
    out = {"smoother": "BJAC/ILU(0) 1 sweep", "solve_s": round(float(np.mean(tm)), 6), "iters": int(it[-1]),
           "relres": er[-1], "vcycle_ms": round(t_vc * 1e3, 3), "setup_s": round(t_b, 2),
           "ilu_levels": levels, "dependent_steps_per_vcycle": steps_vc,
           "us_per_dependent_step": round(t_vc * 1e6 / steps_vc, 3),
           "chain_step_us": round(floor_us, 3),
           "vcycle_lower_bound_ms": round(steps_vc * floor_us * 1e-3, 3),
           "fine_ilu_application_ms": {"grid_substitutions": round(fine_grid_ms, 3),
                                       "level_sets": round(fine_sched_ms, 3)},
           "level_set_steps_levels_1_on": steps_sched,
           "levels_1_on_lower_bound_ms": round(steps_sched * floor_us * 1e-3, 3),
           "note": "dependent steps and the lower bound count every level's level sets; the fine level runs as "
                   "grid substitutions, whose time is fine_ilu_application_ms.grid_substitutions per application"}
    log(f"run B: {steps_vc} dependent steps per V-cycle, {t_vc * 1e6 / steps_vc:.2f} us each; "
        f"one step of a pure chain {floor_us:.2f} us -> lower bound {steps_vc * floor_us * 1e-3:.1f} ms")
    return out


def trsv_chain_step(eng, torch, n=1 << 16, reps=3):
    """Time of one dependent step of the engine's triangular substitution:
    one-level block Jacobi ILU(0) on the 1-D Laplacian, whose factors are
    bidiagonal (every level set one row), so an application is 2n dependent
    steps with nothing beside them."""
    import scipy.sparse as sp

    a = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    p = eng.MldPrec(device=0).init("BJAC")
    p.set("SUB_SOLVE", "ILU")
    p.build(a)
    assert p.tri_levels(0) == (n, n)
    t = time_vcycle(torch, p, n, reps=reps)
    del p
    return t * 1e6 / (2 * n)


def section_diffusion(eng, torch, problems, idim, steps):
    """Heterogeneous diffusion: FP64 CG vs FP32-mixed FCG, dense LU coarse."""
    t0 = time.perf_counter()
    irp, ja, val, bh = problems.diffusion3d(idim, seed=2024)
    t_gen = time.perf_counter() - t0
    n = len(irp) - 1
    b = torch.from_numpy(bh).to("cuda:0")
    p = make_prec(eng, "L1-JACOBI", 2, "double", mcs=4096, coarse="DENSE_LU")
    t0 = time.perf_counter()
    p.hierarchy_build((irp, ja, val))
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    p.smoothers_build()
    t_s = time.perf_counter() - t0
    nc = p.level_info(p.nlev - 1)["n"]
    log(f"diffusion: nlev={p.nlev} coarsest n={nc} hierarchy {t_h:.1f}s smoothers+upload {t_s:.1f}s")
    xd, td, itd, erd = solve_timed(eng, torch, p, "CG", b, 1e-8, steps, 1)
    # FP64 reference solution (same FP64 path, tight tolerance)
    xr = torch.zeros_like(b)
    _, itr, err_r = eng.psb_krylov("CG", None, p, b, xr, eps=1e-12, itmax=2000)
    xref = xr.cpu().numpy()
    e64 = float(np.linalg.norm(xd.cpu().numpy() - xref) / np.linalg.norm(xref))
    p.set_precision("mixed")
    t0 = time.perf_counter()
    p.smoothers_build()
    t_sm = time.perf_counter() - t0
    xm, tmx, itm, erm = solve_timed(eng, torch, p, "FCG", b, 1e-8, steps, 1)
    em = float(np.linalg.norm(xm.cpu().numpy() - xref) / np.linalg.norm(xref))
    rtrue = float(np.linalg.norm(bh - problems.csr_matvec(irp, ja, val, xm.cpu().numpy())) / np.linalg.norm(bh))
    out = {"workload": f"3-D FV diffusion idim={idim}, k=exp(2 g), g ~ N(0,1) (Philox seed 2024) smoothed by 3 "
                       "box-filter passes (edge replication), harmonic faces, b ~ U[-1,1]; L1-JACOBI 2/2, "
                       f"MIN_COARSE_SIZE 4096, coarsest dense FP64 LU (n={nc}, MFMA)",
           "fp64_cg": {"solve_s": round(float(np.mean(td)), 6), "iters": int(itd[-1]), "relres": erd[-1],
                       "rel_err_vs_ref": e64},
           "mixed_fcg": {"solve_s": round(float(np.mean(tmx)), 6), "iters": int(itm[-1]), "relres": erm[-1],
                         "true_relres_host": rtrue, "rel_err_vs_ref": em},
           "speedup_mixed_vs_fp64": round(float(np.mean(td)) / float(np.mean(tmx)), 3),
           "reference_solution": f"FP64 GPU CG to 1e-12 ({itr} iterations, relres {err_r:.2e})",
           "setup_s": {"generate": round(t_gen, 1), "hierarchy_build": round(t_h, 2),
                       "smoothers_build_upload_fp64": round(t_s, 2), "smoothers_build_upload_fp32": round(t_sm, 2)}}
    del p
    return out


def section_block(eng, torch, problems, perfmodel, idim, nrhs, steps):
    """32 seeded Gaussian right-hand sides through the block V-cycle (FP32
    preconditioner, FP64 CG per column) vs one-at-a-time solves."""
    irp, ja, val = problems.pde3d_poisson(idim)
    n = len(irp) - 1
    p = make_prec(eng, "L1-JACOBI", 2, "mixed")
    p.build((irp, ja, val))
    rng = np.random.Generator(np.random.Philox(192))
    B = torch.from_numpy(rng.standard_normal((nrhs, n))).to("cuda:0")
    X = torch.zeros_like(B)
    p.krylov_block("CG", B, X, eps=1e-8)  # warm-up (uploads the block engine, captures nothing)
    tms = []
    for _ in range(steps):
        X.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, it, er = p.krylov_block("CG", B, X, eps=1e-8)
        torch.cuda.synchronize()
        tms.append(time.perf_counter() - t0)
    t_block = float(np.mean(tms))
    # one-at-a-time: the single-vector mixed CG on two of the columns
    singles = []
    its1 = []
    for j in (0, 1):
        xs = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        eng.psb_krylov("CG", None, p, B[j].contiguous(), xs, eps=1e-8)
        torch.cuda.synchronize()
        xs.zero_()
        t0 = time.perf_counter()
        _, i1, _ = eng.psb_krylov("CG", None, p, B[j].contiguous(), xs, eps=1e-8)
        torch.cuda.synchronize()
        singles.append(time.perf_counter() - t0)
        its1.append(i1)
    t_single = float(np.mean(singles)) * nrhs
    Xh = X.cpu().numpy()
    Bh = B.cpu().numpy()
    worst = max(float(np.linalg.norm(Bh[j] - problems.csr_matvec(irp, ja, val, Xh[j])) / np.linalg.norm(Bh[j]))
                for j in range(0, nrhs, 8))
    # block V-cycle throughput
    Xa = torch.ones_like(B)
    p.apply_block(Xa)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    e0.record(s)
    for _ in range(5):
        p.apply_block(Xa)
    e1.record(s)
    e1.synchronize()
    t_bvc = e0.elapsed_time(e1) / 5 / 1e3
    levels = level_table(p)
    vb = perfmodel.block_vcycle_bytes(levels, 4, 4, nrhs, levels[-1]["n"])
    out = {"workload": f"3-D Poisson idim={idim} (n={n}), {nrhs} RHS ~ N(0,1) (Philox seed 192), FP32 preconditioner "
                       "(L1-JACOBI 2/2, coarsest ILU(0) as explicit operator), FP64 CG per column to 1e-8",
           "block_solve_s": round(t_block, 5), "iters_max": int(it.max()), "iters_min": int(it.min()),
           "relres_max": float(er.max()), "true_relres_max_sampled": worst,
           "one_at_a_time_s": round(t_single, 5), "one_at_a_time_iters": its1,
           "speedup_vs_one_at_a_time": round(t_single / t_block, 3),
           "block_vcycle_ms": round(t_bvc * 1e3, 3), "block_vcycle_GBps": round(vb / t_bvc / 1e9, 1)}
    del p
    return out


def section_group(eng, torch, problems, nsys, lo, hi, steps):
    """64 independent Poisson systems, idim uniform in [lo, hi], one RHS each,
    solved together (one union V-cycle, iterations in one hipGraph) vs one
    system at a time."""
    rng = np.random.Generator(np.random.Philox(64))
    idims = rng.integers(lo, hi + 1, size=nsys)
    systems = []
    t0 = time.perf_counter()
    for d in idims:
        p = make_prec(eng, "L1-JACOBI", 2, "mixed")
        p.build(problems.pde3d_poisson(int(d)))
        systems.append(p)
    t_b = time.perf_counter() - t0
    g = eng.MldGroup(device=0, nb=1)
    for p in systems:
        g.add(p)
    t0 = time.perf_counter()
    g.build("mixed")
    t_g = time.perf_counter() - t0
    ns, nr, off = g.info()
    bh = rng.standard_normal(nr)
    b = torch.from_numpy(bh).to("cuda:0")
    x, it, er, nl = g.krylov("FCG", b, eps=1e-8)  # warm-up + graph capture
    tms = []
    for _ in range(steps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, it, er, nl = g.krylov("FCG", b, x, eps=1e-8)
        torch.cuda.synchronize()
        tms.append(time.perf_counter() - t0)
    t_group = float(np.mean(tms))
    # one system at a time (same handles, single-system device path)
    t_seq = 0.0
    it_seq = []
    for s, p in enumerate(systems):
        bs = b[off[s]:off[s + 1]].contiguous()
        xs = torch.zeros_like(bs)
        eng.psb_krylov("FCG", None, p, bs, xs, eps=1e-8)
        xs.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, i1, _ = eng.psb_krylov("FCG", None, p, bs, xs, eps=1e-8)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        it_seq.append(i1)
    t_apply = g.time_apply(5)
    out = {"workload": f"{nsys} 3-D Poisson systems, idim uniform in [{lo},{hi}] (Philox seed 64; total n={nr}), "
                       "RHS ~ N(0,1), FP32 preconditioners (L1-JACOBI 2/2, coarsest ILU(0) explicit), FP64 FCG per "
                       "system to 1e-8",
           "idims": [int(d) for d in idims], "nlev_range": [min(p.nlev for p in systems), max(p.nlev for p in systems)],
           "group_solve_s": round(t_group, 5), "graph_launches": nl, "iters_max": int(it.max()),
           "relres_max": float(er.max()), "one_at_a_time_s": round(t_seq, 5), "one_at_a_time_iters_max": max(it_seq),
           "speedup_vs_one_at_a_time": round(t_seq / t_group, 3), "group_vcycle_ms": round(t_apply, 3),
           "setup_s": {"hierarchies": round(t_b, 2), "group_union_upload": round(t_g, 2)}}
    del g, systems
    return out


def section_adjoint(eng, torch, problems, perfmodel, idim):
    """dL/dtheta at idim: forward mixed CG to 1e-10, adjoint solve with the
    same hierarchy, edge-product gradient kernel."""
    from mld2p4_amd import autograd as ag

    th, bh, ths = problems.theta_problem(idim, seed=5)
    opts = ag.SolveOptions(method="CG", eps=1e-10, precision="mixed")
    # target x* at the second field (same solver)
    with torch.no_grad():
        xs = ag.solve(torch.from_numpy(ths).to("cuda:0"), torch.from_numpy(bh).to("cuda:0"), idim, opts)
    t = torch.from_numpy(th).to("cuda:0").requires_grad_(True)
    b = torch.from_numpy(bh).to("cuda:0").requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = ag.solve(t, b, idim, opts)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    loss = 0.5 * torch.sum((x - xs) ** 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    last = dict(opts.last)
    # gradient kernel alone, HIP events on torch's stream (the kernel's stream)
    lam = torch.randn_like(x)
    gbuf = torch.empty_like(x)
    wbuf = torch.empty_like(x)
    th_d = t.detach()
    eng.fv7_grad_theta(idim, th_d, x.detach(), lam, gbuf, wbuf)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # timing loop on the raw C entry point with prepared arguments (the
    # Python wrapper's checks would make the loop host-bound)
    import ctypes as C

    fn = eng.lib().mldhip_fv7_grad_theta
    sp = C.c_void_p(s.cuda_stream)
    args = [C.c_void_p(t_.data_ptr()) for t_ in (th_d, x.detach(), lam, gbuf)]

    def loop(work):
        w = C.c_void_p(work.data_ptr()) if work is not None else None
        for _ in range(5):
            fn(idim, *args, w, sp)
        e0.record(s)
        for _ in range(50):
            fn(idim, *args, w, sp)
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / 50 / 1e3

    t_ws = loop(wbuf)
    t_k1 = loop(None)
    t_k = min(t_ws, t_k1)
    by = perfmodel.fv7_grad_bytes(idim ** 3)
    return {"workload": f"3-D FV diffusion idim={idim}, theta ~ N(0, 0.5^2) (Philox seed 5), b ~ U[-1,1], x* at "
                        "theta* (seed 6); L = ||x - x*||^2/2; FP32-mixed AMG-CG to 1e-10 forward, adjoint CG with the "
                        "same hierarchy transposed",
            "forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4), "solves": last,
            "loss": float(loss.detach()), "grad_theta_norm": float(t.grad.norm()),
            "grad_kernel": {"avg_ms": round(t_k * 1e3, 4), "algorithmic_bytes": by,
                            "variants_ms": {"exp_workspace": round(t_ws * 1e3, 4), "exp_in_place": round(t_k1 * 1e3, 4)},
                            "GBps": round(by / t_k / 1e9, 1), "frac_of_hbm_peak": round(by / t_k / 1e9 / HBM_PEAK_GBS, 4)},
            "note": "setup (host assembly + hierarchy build + upload) is inside forward_s"}


def traffic_probe(idim, tuned=False):
    """Launch sequence for counter runs: 3 copies of 2^27 doubles (1 GiB read
    + 1 GiB written each), then 3 launches of each fine-level step of the FP64
    benchmark hierarchy and of level 1's residual and transfers.  Storage
    autotuning is off unless `tuned` (the fine level is DIA regardless; with
    tuning the level-1 layouts are the benchmark's, and the probe's launches
    are the last ones of each kernel in the trace)."""
    import torch

    if not tuned:
        os.environ["MLD_AUTOTUNE"] = "0"
    eng = _engine()
    from mld2p4_amd import problems

    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    irp, ja, val = problems.pde3d_poisson(idim)
    p = make_prec(eng, "L1-JACOBI", 2, "double")
    p.build((irp, ja, val))
    for kind in ("residual", "pre", "post", "krylov_spmv"):
        p.time_step(0, kind, 2)  # one warm-up launch sequence + 2 timed
    for kind in ("restrict", "prolong", "residual"):
        p.time_step(1, kind, 2)
    torch.cuda.synchronize()
    log("traffic probe done")


def block_probe(idim=192, nrhs=32):
    """Launch sequence for counter runs of the block V-cycle: the same three
    1 GiB copies (calibration), then 3 block applications (FP32 preconditioner,
    L1-JACOBI 2/2, storage autotuning on) of nrhs right-hand sides at idim^3;
    the last application's kernels are the measured ones."""
    import torch

    eng = _engine()
    from mld2p4_amd import problems

    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    irp, ja, val = problems.pde3d_poisson(idim)
    p = make_prec(eng, "L1-JACOBI", 2, "mixed")
    p.build((irp, ja, val))
    X = torch.ones((nrhs, len(irp) - 1), dtype=torch.float64, device="cuda:0")
    for _ in range(3):
        p.apply_block(X)
    torch.cuda.synchronize()
    log("block probe done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--idim", type=int, default=256)
    ap.add_argument("--smoother", default="L1-JACOBI")
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-sample-iters", type=int, default=5)
    ap.add_argument("--sections", default="c1,run_b,diffusion,batched,adjoint",
                    help="extra sections to run (comma list; empty for none)")
    ap.add_argument("--small", action="store_true", help="small sizes for the extra sections (quick check)")
    ap.add_argument("--tuned", action="store_true", help="traffic probe with storage autotuning on")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the kernels whose HBM traffic is counted under rocprofv3 --pmc: a 1 GiB copy "
                         "(the calibration kernel), then the fine-level steps of the benchmark's hierarchy")
    ap.add_argument("--block-probe", action="store_true",
                    help="only run the kernels of block V-cycles (192^3, 32 right-hand sides) for rocprofv3 --pmc, "
                         "after the same 1 GiB calibration copies")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box here has one GPU). "
              "How the path would shard is a design only.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args.idim, tuned=args.tuned)
        return
    if args.block_probe:
        block_probe()
        return

    import torch

    eng = _engine()
    from mld2p4_amd import problems, perfmodel  # noqa: E402  (package registered by _engine)

    idim = args.idim
    sections = [s for s in args.sections.split(",") if s]
    log(f"generating 3-D Poisson idim={idim}")
    t0 = time.perf_counter()
    irp, ja, val = problems.pde3d_poisson(idim)
    n = len(irp) - 1
    log(f"n={n} nnz={len(ja)} generated in {time.perf_counter() - t0:.1f}s")

    # ---------------- FP64 path
    t0 = time.perf_counter()
    p = make_prec(eng, args.smoother, args.sweeps, "double")
    p.hierarchy_build((irp, ja, val))
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    p.smoothers_build()
    t_s = time.perf_counter() - t0
    log(f"hierarchy_build {t_h:.1f}s smoothers_build+upload {t_s:.1f}s nlev={p.nlev} "
        f"complexity={p.complexity():.3f}")
    levels = level_table(p)

    def fmt_str(l, what):
        f = p.op_format(l, what)
        return f["fmt"] + (f"x{f['lanes']}" if f["fmt"] == "CSR" else "")

    for l, L in enumerate(levels):
        ops = " ".join(f"{w}:{fmt_str(l, w)}" for w in (["A", "ND"] if l == 0 else ["A", "ND", "P", "R"])
                       if not (w == "ND" and L["nnz_nd"] == 0))
        log(f"  level {l}: n={L['n']} nnz={L['nnz_a']} nnzP={L['nnz_p']} {ops}")
    b = torch.ones(n, dtype=torch.float64, device="cuda:0")
    log("FP64 CG solves")
    x, times, its, errs = solve_timed(eng, torch, p, "CG", b, 1e-8, args.steps, args.warmup)
    t_solve = float(np.mean(times))
    iters = int(its[-1])
    log(f"FP64 CG: {t_solve * 1e3:.2f} ms/solve, iters={its}, relres={errs[-1]:.3e}")
    xh = x.cpu().numpy()
    rtrue = np.linalg.norm(1.0 - problems.csr_matvec(irp, ja, val, xh)) / np.sqrt(n)

    # V-cycle throughput
    t_vc = time_vcycle(torch, p, n)
    vbytes, vsteps = perfmodel.vcycle_bytes(levels, 8, 8)
    log(f"V-cycle {t_vc * 1e3:.3f} ms, algorithmic {vbytes / 1e9:.2f} GB -> {vbytes / t_vc / 1e9:.0f} GB/s")

    # dominant kernel: fine-level steps timed with HIP events on the handle stream
    L0, L1 = levels[0], levels[1]
    cand = {}
    reps = 20
    cand["residual"] = (p.time_step(0, "residual", reps),
                        perfmodel.residual_bytes(n, L0["nnz_a"], 8, 8, fmt=L0["fmt_a"]), 2)
    cand["pre"] = (p.time_step(0, "pre", reps),
                   perfmodel.smoother_bytes(n, L0["solver"], L0["sweeps_pre"], L0["nnz_nd"], 0, 8, 8, 0,
                                            nd_fmt=L0.get("fmt_nd")), 1)
    cand["post"] = (p.time_step(0, "post", reps),
                    perfmodel.smoother_bytes(n, L0["solver"], L0["sweeps_post"], L0["nnz_nd"], 0, 8, 8, 1,
                                             nd_fmt=L0.get("fmt_nd")), 1)
    cand["restrict"] = (p.time_step(1, "restrict", reps),
                        perfmodel.spmv_bytes(L1["n"], n, L1["nnz_r"], 8, 8, fmt=L1["fmt_r"]), 1)
    cand["prolong"] = (p.time_step(1, "prolong", reps),
                       perfmodel.spmv_bytes(n, L1["n"], L1["nnz_p"], 8, 8, beta_read=True, fmt=L1["fmt_p"]), 1)
    cand["krylov_spmv"] = (p.time_step(0, "krylov_spmv", reps),
                           perfmodel.spmv_bytes(n, n, L0["nnz_a"], 8, 8, fmt=L0["fmt_a"]), 1)
    for k, (ms, by, cnt) in cand.items():
        log(f"  step {k}: {ms * 1e3:.1f} us, {by / 1e6:.1f} MB -> {by / ms / 1e6:.0f} GB/s")
    dom = max(cand, key=lambda k: cand[k][0] * cand[k][2])
    ms, by, _ = cand[dom]
    if dom == "residual":
        # the dominant kernel's figure: its launches inside whole V-cycles
        ms = p.time_step(0, "residual_in_cycle", reps)
        log(f"  residual inside V-cycles: {ms * 1e3:.1f} us -> {by / ms / 1e6:.0f} GB/s")
    achieved = by / (ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(dom)
        except Exception:
            traffic = None
    copy_gbs = copy_bandwidth(torch)
    log(f"device-to-device copy of 2 GiB (HIP runtime copy kernel): {copy_gbs:.0f} GB/s")
    roofline = {"bound": "hbm", "kernel": f"fine-level {dom} (level 1 of {p.nlev}), launches inside V-cycles",
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": int(by), "avg_ms": round(ms, 5),
                "steps_GBps": {k: round(v[1] / v[0] / 1e6, 1) for k, v in cand.items()},
                "hip_memcpy_d2d_GBps": round(copy_gbs, 1)}

    # ---------------- FP32-mixed path (FP32 hierarchy inside FP64 FCG)
    mixed = None
    if not args.no_mixed:
        # same host hierarchy, uploaded again with FP32 values
        p.set_precision("mixed")
        p.smoothers_build()
        torch.cuda.synchronize()
        log("mixed FCG solves")
        xm, tm, itm, erm = solve_timed(eng, torch, p, "FCG", b, 1e-8, args.steps, args.warmup)
        t_vcm = time_vcycle(torch, p, n)
        vbm, _ = perfmodel.vcycle_bytes(levels, 4, 4)
        xd = xm.cpu().numpy()
        mixed = {"solve_s": round(float(np.mean(tm)), 6), "iters": int(itm[-1]), "relres": erm[-1],
                 "speedup_vs_fp64": round(t_solve / float(np.mean(tm)), 3),
                 "rel_err_vs_fp64_solution": float(np.linalg.norm(xd - xh) / np.linalg.norm(xh)),
                 "vcycle_ms": round(t_vcm * 1e3, 4), "vcycle_GBps": round(vbm / t_vcm / 1e9, 1)}
        log(f"mixed FCG: {mixed}")
    cplx = p_cplx(levels)
    del p
    release(torch)

    extra = {}
    small = args.small
    if "c1" in sections:
        log("C1: idim 40, oracle on one core vs GPU")
        extra["c1"] = section_c1(eng, torch, problems)
        log(f"C1: {extra['c1']}")
        release(torch)
    if "run_b" in sections:
        log("run B: BJAC/ILU(0) smoother")
        extra["run_b"] = section_run_b(eng, torch, problems, irp, ja, val, n, min(args.steps, 2))
        log(f"run B: {extra['run_b']}")
        release(torch)
    del irp, ja, val
    if "diffusion" in sections:
        log("diffusion: FP64 vs mixed")
        extra["diffusion"] = section_diffusion(eng, torch, problems, 48 if small else idim, min(args.steps, 3))
        log(f"diffusion: {extra['diffusion']}")
        release(torch)
    if "batched" in sections:
        log("batched: 32 RHS block V-cycle")
        extra["batched_block"] = section_block(eng, torch, problems, perfmodel, 40 if small else 192, 32,
                                               min(args.steps, 3))
        log(f"batched block: {extra['batched_block']}")
        release(torch)
        log("batched: 64 grouped systems")
        extra["batched_group"] = section_group(eng, torch, problems, 64, 8 if small else 24, 24 if small else 96,
                                               min(args.steps, 3))
        log(f"batched group: {extra['batched_group']}")
        release(torch)
    if "adjoint" in sections:
        log("adjoint: dL/dtheta")
        extra["adjoint"] = section_adjoint(eng, torch, problems, perfmodel, 32 if small else 128)
        log(f"adjoint: {extra['adjoint']}")
        release(torch)

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle) ...")
        cpu = cpu_baseline(idim, iters, args.cpu_sample_iters)
        log(f"CPU baseline: {cpu}")

    out = {
        "metric": "AMG-PCG time to rel.res 1e-8 (s) + V-cycle HBM GB/s; FP32-mixed vs FP64",
        "value": round(t_solve, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_solve * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Poisson 7-point idim={idim} (n={n}), FP64 values, int32 indices, "
                               f"b=1, x0=0; SA-AMG V-cycle {args.sweeps}/{args.sweeps} {args.smoother} sweeps, "
                               "coarsest BJAC/ILU(0), CG to 1e-8",
                   "idim": idim, "nlev": len(levels), "operator_complexity": round(cplx, 4)},
        "iterations": iters,
        "relres": errs[-1],
        "true_relres_host": float(rtrue / 1.0),
        "vcycle_ms": round(t_vc * 1e3, 4),
        "vcycles_per_s": round(1.0 / t_vc, 1),
        "vcycle_GBps": round(vbytes / t_vc / 1e9, 1),
        "setup_s": {"hierarchy_build": round(t_h, 2), "smoothers_build_upload": round(t_s, 2)},
        "roofline": roofline,
        "mixed": mixed,
        **extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def p_cplx(levels):
    return sum(L["nnz_a"] for L in levels) / levels[0]["nnz_a"]


if __name__ == "__main__":
    main()
