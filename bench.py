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
per solve over the timed steps.  The FP32-mixed form (FP32 hierarchy inside
FP64 FCG) of the same problem is reported beside it.
"""
from __future__ import annotations

import argparse
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


def make_prec(eng, A, smoother, sweeps, precision):
    p = eng.MldPrec(device=0).init("ML")
    p.set("SMOOTHER_TYPE", smoother)
    p.set("SMOOTHER_SWEEPS", sweeps)
    p.set("AGGR_THRESH", 0.0)
    p.set("AGGR_PROL", "SMOOTHED")
    p.set("MIN_COARSE_SIZE", 200)
    p.set("COARSE_SOLVE", "BJAC")
    p.set("COARSE_SUBSOLVE", "ILU")
    p.set("COARSE_SWEEPS", 4)
    p.set_precision(precision)
    return p


def level_table(eng, p):
    from importlib import import_module

    levels = []
    for l in range(p.nlev):
        i = p.level_info(l)
        d = dict(n=i["n"], nnz_a=i["nnz_a"], nnz_p=i["nnz_p"], nnz_r=i["nnz_r"], nnz_nd=i["nnz_nd"],
                 solver=i["solver"], sweeps_pre=i["sweeps_pre"], sweeps_post=i["sweeps_post"])
        fa = p.op_format(l, "A")
        d["csr_a"] = fa["fmt"] == "CSR"
        if i["nnz_nd"] > 0:
            d["csr_nd"] = p.op_format(l, "ND")["fmt"] == "CSR"
        if i["solver"] == 7:
            d["nnz_lu"] = p.level_csr(l, "L").nnz + p.level_csr(l, "U").nnz
        levels.append(d)
    return levels


def solve_timed(eng, torch, A, p, method, b, eps, steps, warmup):
    x = torch.zeros_like(b)
    for _ in range(warmup):
        x.zero_()
        eng.psb_krylov(method, None, p, b, x, eps=eps, itmax=1000)
    torch.cuda.synchronize()
    times, its, errs = [], [], []
    for _ in range(steps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, it, err = eng.psb_krylov(method, None, p, b, x, eps=eps, itmax=1000)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        its.append(it)
        errs.append(err)
    return x, times, its, errs


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


def cpu_baseline(idim, gpu_iters, sample_iters):
    """Oracle (sequential CPU restatement) timed on this host: hierarchy
    build, then a bounded sample of CG iterations on the same problem, scaled
    to the GPU's iteration count."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O

    a = O.gen_pde3d(idim)
    prm = O.config1_params(smoother=O.SM_L1JAC, sweeps_pre=2, sweeps_post=2)
    t0 = time.perf_counter()
    po = O.Prec(prm).build(a)
    tb = time.perf_counter() - t0
    b = np.ones(a.shape[0])
    t0 = time.perf_counter()
    _, it, _ = po.krylov("CG", b, eps=1e-30, itmax=sample_iters)
    ts = time.perf_counter() - t0
    per_it = ts / max(it, 1)
    return {
        "value": per_it * gpu_iters,
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (sequential C++ restatement) at idim={idim}: hierarchy build {tb:.1f}s (not in value), "
                   f"{it} CG iterations timed ({ts:.1f}s, {per_it:.2f}s/iter) x {gpu_iters} GPU iterations"),
    }


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
    ap.add_argument("--cpu-sample-iters", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box here has one GPU). "
              "The multi-GPU sharding design is described in DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    import torch

    eng = _engine()
    from mld2p4_amd import problems, perfmodel  # noqa: E402  (package registered by _engine)

    idim = args.idim
    log(f"generating 3-D Poisson idim={idim}")
    t0 = time.perf_counter()
    irp, ja, val = problems.pde3d_poisson(idim)
    n = len(irp) - 1
    log(f"n={n} nnz={len(ja)} generated in {time.perf_counter() - t0:.1f}s")

    # ---------------- FP64 path
    t0 = time.perf_counter()
    p = make_prec(eng, None, args.smoother, args.sweeps, "double")
    p.hierarchy_build((irp, ja, val))
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    p.smoothers_build()
    t_s = time.perf_counter() - t0
    log(f"hierarchy_build {t_h:.1f}s smoothers_build+upload {t_s:.1f}s nlev={p.nlev} "
        f"complexity={p.complexity():.3f}")
    levels = level_table(eng, p)
    def fmt_str(l, what):
        f = p.op_format(l, what)
        return f["fmt"] + (f"x{f['lanes']}" if f["fmt"] == "CSR" else "")

    for l, L in enumerate(levels):
        ops = " ".join(f"{w}:{fmt_str(l, w)}" for w in (["A", "ND"] if l == 0 else ["A", "ND", "P", "R"])
                       if not (w == "ND" and L["nnz_nd"] == 0))
        log(f"  level {l}: n={L['n']} nnz={L['nnz_a']} nnzP={L['nnz_p']} {ops}")
    b = torch.ones(n, dtype=torch.float64, device="cuda:0")
    log("FP64 CG solves")
    x, times, its, errs = solve_timed(eng, torch, None, p, "CG", b, 1e-8, args.steps, args.warmup)
    t_solve = float(np.mean(times))
    iters = int(its[-1])
    log(f"FP64 CG: {t_solve * 1e3:.2f} ms/solve, iters={its}, relres={errs[-1]:.3e}")
    # true residual on the host (FP64)
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
                        perfmodel.residual_bytes(n, L0["nnz_a"], 8, 8, L0["csr_a"]), 2)
    cand["pre"] = (p.time_step(0, "pre", reps),
                   perfmodel.smoother_bytes(n, L0["solver"], L0["sweeps_pre"], L0["nnz_nd"], 0, 8, 8, 0,
                                            L0.get("csr_nd", False)), 1)
    cand["post"] = (p.time_step(0, "post", reps),
                    perfmodel.smoother_bytes(n, L0["solver"], L0["sweeps_post"], L0["nnz_nd"], 0, 8, 8, 1,
                                             L0.get("csr_nd", False)), 1)
    cand["restrict"] = (p.time_step(1, "restrict", reps), perfmodel.spmv_bytes(L1["n"], n, L1["nnz_r"], 8, 8), 1)
    cand["prolong"] = (p.time_step(1, "prolong", reps),
                       perfmodel.spmv_bytes(n, L1["n"], L1["nnz_p"], 8, 8, beta_read=True), 1)
    cand["krylov_spmv"] = (p.time_step(0, "krylov_spmv", reps),
                           perfmodel.spmv_bytes(n, n, L0["nnz_a"], 8, 8, csr=L0["csr_a"]), 1)
    for k, (ms, by, cnt) in cand.items():
        log(f"  step {k}: {ms * 1e3:.1f} us, {by / 1e6:.1f} MB -> {by / ms / 1e6:.0f} GB/s")
    dom = max(cand, key=lambda k: cand[k][0] * cand[k][2])
    ms, by, _ = cand[dom]
    achieved = by / (ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(dom)
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "kernel": f"fine-level {dom} (level 1 of {p.nlev})", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": int(by), "avg_ms": round(ms, 5)}

    # ---------------- FP32-mixed path (FP32 hierarchy inside FP64 FCG)
    mixed = None
    if not args.no_mixed:
        # same host hierarchy, uploaded again with FP32 values
        pm = p
        pm.set_precision("mixed")
        pm.smoothers_build()
        torch.cuda.synchronize()
        log("mixed FCG solves")
        xm, tm, itm, erm = solve_timed(eng, torch, None, pm, "FCG", b, 1e-8, args.steps, args.warmup)
        t_vcm = time_vcycle(torch, pm, n)
        vbm, _ = perfmodel.vcycle_bytes(levels, 4, 4)
        xd = xm.cpu().numpy()
        mixed = {"solve_s": round(float(np.mean(tm)), 6), "iters": int(itm[-1]), "relres": erm[-1],
                 "speedup_vs_fp64": round(t_solve / float(np.mean(tm)), 3),
                 "rel_err_vs_fp64_solution": float(np.linalg.norm(xd - xh) / np.linalg.norm(xh)),
                 "vcycle_ms": round(t_vcm * 1e3, 4), "vcycle_GBps": round(vbm / t_vcm / 1e9, 1)}
        log(f"mixed FCG: {mixed}")
    del p

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
        "config": {"workload": f"3-D Poisson 7-point idim={idim} (n={n}, nnz={len(ja)}), FP64 values, int32 indices, "
                               f"b=1, x0=0; SA-AMG V-cycle {args.sweeps}/{args.sweeps} {args.smoother} sweeps, "
                               "coarsest BJAC/ILU(0), CG to 1e-8",
                   "idim": idim, "nlev": len(levels), "operator_complexity": round(p_cplx(levels), 4)},
        "iterations": iters,
        "relres": errs[-1],
        "true_relres_host": float(rtrue / 1.0),
        "vcycle_ms": round(t_vc * 1e3, 4),
        "vcycles_per_s": round(1.0 / t_vc, 1),
        "vcycle_GBps": round(vbytes / t_vc / 1e9, 1),
        "setup_s": {"hierarchy_build": round(t_h, 2), "smoothers_build_upload": round(t_s, 2)},
        "roofline": roofline,
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def p_cplx(levels):
    return sum(L["nnz_a"] for L in levels) / levels[0]["nnz_a"]


if __name__ == "__main__":
    main()
