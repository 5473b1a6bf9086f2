#!/usr/bin/env python3
"""Benchmark of the MI355X SIMP state-solve path.

Headline: SIMP iterations/s on the 160x80x80 hex8 cantilever (1,024,000
elements, 3.17M DOF; x=0 face clamped, total -1 z line load on the x=160,
z=0 edge; volfrac 0.3, rmin 1.5, move 0.2; Galerkin GMG-PCG with
Chebyshev-Jacobi smoothing; PCG tol 1e-5 relative residual, max_krylov 800),
for the BF16-MFMA path and the plain FP32 path. A step is one full SIMP
iteration: filter, projection, moduli, GMG setup, state solve to tol (FP64
true residual), compliance, sensitivity, filter transpose, OC bisection.
Inputs are resident in HBM; the density starts at volfrac, W warm-up
iterations are run untimed, then K iterations are timed between device
synchronisations (defaults 2 + 48: the workload's 50 OC iterations). The
hierarchy is the solver's default (coarsened until the coarsest level fits
the one-workgroup dense inverse; Chebyshev degree 1 / 1 / 2 / 10 / 4 on level
0 / 1 / 2 / 3 / below); the solver of the workload description (4 levels,
degree 3)
is reported as a variant.

Also reported: matvec GB/s on the 128x64x64 grid (524,288 elements,
algorithmic bytes 8*ndof + 4*n_elem per product); the roofline of the
kernel with the most device time per step among every operator and grid
transfer launch of the V-cycle and the PCG (all levels), with the coarse
levels' dominant kernel beside it; and a CPU baseline: the oracle's
run_simp (FP64, spsolve, one thread) timed on the 24x12x6 cantilever next to
the GPU on the same case, plus a labelled model of the 1M-element step.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

METRIC = "SIMP iter/s @1M hex8 (BF16-MFMA GMG vs FP32 plain, rtol 1e-5); matvec GB/s"
# The 1M-element cantilever of the headline. The hierarchy is the solver's
# default: coarsened until the coarsest level fits one workgroup's dense
# inverse (6 levels here), Chebyshev degree 1 on levels 0 and 1, 2 on level
# 2, 10 on level 3 and 4 below (the expensive levels smooth least, the
# launch-bound small ones most; tuned on this workload, which the workload
# description's 4 levels / degree 3 reports as a variant).
C2 = dict(dims=(160, 80, 80), volfrac=0.30, rmin=1.5, move=0.2, n_levels="auto", fine_degree=1,
          coarse_degree=(1, 2, 10, 4), tol=1e-5, max_krylov=800)
C1 = (128, 64, 64)
# Multi-load bracket: 192x64x64, y=0 face clamped, 8 unit point loads on the
# y=64 face (nodes and directions from default_rng(3)), equal weights, the 8
# load cases solved as one block of right-hand sides; 30 OC iterations.
C4 = dict(dims=(192, 64, 64), n_loads=8, seed=3, volfrac=0.30, rmin=1.5, move=0.2, n_levels="auto", fine_degree=1,
          coarse_degree=(1, 2, 10, 4), tol=1e-5, max_krylov=800)
HBM_PEAK_GBPS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0


def degree_schedule(cfg: dict) -> str:
    """'1 / 1 / 2 / 10 / 4 on level 0 / 1 / 2 / 3 / below' from a case's
    fine and coarse degrees (the last coarse entry repeats below)."""
    degs = [cfg["fine_degree"], *cfg["coarse_degree"]]
    levels = [str(k) for k in range(len(degs) - 1)] + ["below"]
    return f"{' / '.join(map(str, degs))} on level {' / '.join(levels)}"


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# HBM bytes and MFMA busy fraction per launch of the probed (level, kernel,
# epilogue) launches at 160x80x80, measured with hardware counters
# (FETCH_SIZE / WRITE_SIZE runs of ``bench.py --kernel-probe`` scaled by the
# counts of a 256 MiB copy, and SQ_VALU_MFMA_BUSY_CYCLES); the summary file
# holds the commands and the figures, one "<field>[<probe name>]: <value>"
# line per kernel.
COUNTER_PROFILE = "profiles/counters_kernel_probe_160x80x80.md"
PROBE_NAMES = {"L0 hex_bf16_mfma/CHEB_R": "L0 CHEB_R", "L0 hex_bf16_mfma/CHEB_FIRST": "L0 CHEB_FIRST",
               "L0 hex_bf16_mfma/RESID_UPDATE": "L0 RESID_UPDATE", "L0 hex_bf16_mfma/RESID_FIRST": "L0 RESID_FIRST",
               "L0 hex_bf16_mfma/CHEB_LAST": "L0 CHEB_LAST",
               "L0 hex_bf16x3_mfma/DOT": "L0 split DOT", "L0 hex_bf16_mfma/STORE": "L0 STORE",
               "L1 galerkin_q8_bf16_mfma/RESID_FIRST": "L1 RESID_FIRST",
               "L1 galerkin_q8_bf16_mfma/CHEB_LAST": "L1 CHEB_LAST", "L2 stencil_bf16/CHEB_R": "L2 CHEB_R",
               "L3 stencil_bf16/CHEB_R": "L3 CHEB_R"}


def profiled_value(kernel: str, field: str):
    """A counter figure for the kernel from the counter summary under
    profiles/, as (value, path), or (None, None) when it holds none."""
    probe = PROBE_NAMES.get(kernel)
    if probe is None:
        return None, None
    full = os.path.join(os.path.dirname(os.path.abspath(__file__)), COUNTER_PROFILE)
    tag = f"{field}[{probe}]:"
    try:
        with open(full) as f:
            for line in f:
                if line.startswith(tag):
                    return float(line[len(tag):]), COUNTER_PROFILE
    except OSError:
        pass
    return None, None


def make_engine(case: str, precision: str, n_levels=None, loads=None, **over):
    """HipSIMP for the 1M-element cantilever ("c2") or the multi-load
    bracket ("c4"; ``loads`` selects a subset of its load cases); ``over``
    replaces entries of the case's solver settings."""
    import topogpu_amd as tg
    from topogpu_amd.bc import generate_bc
    from topogpu_amd.hip.simp import HipSIMP

    if case == "c2":
        cfg = C2
        prob = tg.unit_cantilever(cfg["dims"], volfrac=cfg["volfrac"], rmin=cfg["rmin"], load="line")
        force = prob.force_override
    else:
        cfg = C4
        prob = tg.multi_load_bracket(cfg["dims"], n_loads=cfg["n_loads"], seed=cfg["seed"], volfrac=cfg["volfrac"],
                                     rmin=cfg["rmin"])
        force = prob.force_override if loads is None else prob.force_override[list(loads)]
    bc = generate_bc(prob.spec)
    cfg = {**cfg, **over}
    return HipSIMP(cfg["dims"], bc.free_dofs, force, volfrac=cfg["volfrac"], rmin=cfg["rmin"], move=cfg["move"],
                   precision=precision, n_levels=n_levels or cfg["n_levels"], fine_degree=cfg["fine_degree"],
                   coarse_degree=cfg["coarse_degree"], tol=cfg["tol"], max_krylov=cfg["max_krylov"])


def simp_bench(precision: str, steps: int, warmup: int, n_levels=None, saturated: bool = False, case: str = "c2",
               loads=None, **over):
    """Run warmup + steps SIMP iterations from rho = volfrac; time the steps.
    Operator and grid-transfer launches of every level are counted per
    (level, kernel, epilogue) in the timed region (no events: nothing but a
    counter is added to the loop)."""
    import torch

    from topogpu_amd.hip.gmg import EP_NAMES
    from topogpu_amd.hip.timing import KernelTimer

    eng = make_engine(case, precision, n_levels, loads, **over)
    for i in range(warmup):
        r = eng.step()
        _log(f"{precision} warmup {i + 1}/{warmup}: C={r.compliance:.6g} its={r.iters} conv={r.converged}")
    counter = KernelTimer(events=False)
    eng.fea.gmg.timer = counter
    # the long-lived heap (torch, numpy, the engine) moves to the collector's
    # permanent generation, so a full collection in the loop does not rescan
    # it (a 30 ms pause when one happens); every step still runs in full
    gc.collect()
    gc.freeze()
    torch.cuda.synchronize()
    rows = []
    t0 = time.perf_counter()
    for i in range(steps):
        r = eng.step()
        rows.append(r)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    gc.unfreeze()
    eng.fea.gmg.timer = None
    counter.flush()
    for i, r in enumerate(rows):
        if i < 3 or i >= steps - 3 or (i + 1) % 10 == 0:
            _log(f"{precision} step {i + 1}/{steps}: C={r.compliance:.6g} its={r.iters} restarts={r.restarts} "
                 f"relres={r.rel_residual:.2e} oc={r.oc_steps}")
    stats = {
        "iter_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps, "n_levels": len(eng.fea.gmg.levels),
        "coarsest_dofs": eng.fea.gmg.n_coarse,
        "pcg_iters_mean": float(np.mean([r.iters for r in rows])),
        "pcg_iters_max": int(max(r.iters for r in rows)),
        "ir_restarts_mean": float(np.mean([r.restarts for r in rows])),
        "max_rel_residual": float(max(r.rel_residual for r in rows)),
        "all_converged": bool(all(r.converged for r in rows)),
        "compliance_last": float(rows[-1].compliance), "n_loads": eng.fea.nrhs,
    }
    launches = {f"{name}/{EP_NAMES[mode]}": n / steps for (name, mode), n in counter.counts.items()}
    kernels = {}
    if saturated:
        kernels = saturated_kernel_times(counter, steps)
    n = {"ndof": eng.fea.n, "n_elem": eng.fea.n_elem}
    del eng
    torch.cuda.empty_cache()
    return stats, launches, kernels, n


def saturated_kernel_times(counter, steps: int, reps: int = 50) -> dict:
    """Device duration of every accounted (level, kernel, epilogue) launch
    of the timed region: its last launch (live operands of the last SIMP
    iteration) captured reps times back to back into one HIP graph on a side
    stream and replayed between two HIP events recorded on that stream, so
    host launch rates and gaps do not enter (a kernel shorter than the host's
    launch interval would otherwise be timed at that interval). The average
    is the kernel's duration plus the graph's dispatch gap, about 1 us. A
    reducing launch includes its one-block finalize. Run after the timed
    region; the replays overwrite the solver's work vectors."""
    import ctypes

    import torch

    from topogpu_amd.hip import lib as L
    from topogpu_amd.hip.gmg import EP_NAMES

    out = {}
    side = torch.cuda.Stream()
    st = side.cuda_stream
    for (name, mode), (fn, nb) in counter.replay.items():
        torch.cuda.synchronize()
        L.call("tg_graph_begin", st)
        h = ctypes.c_void_p()
        try:
            for _ in range(reps):
                fn(st)
        finally:
            L.load().tg_graph_end(st, ctypes.byref(h))
        try:
            L.call("tg_graph_launch", h.value, st)   # warm-up replay
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(side)
            L.call("tg_graph_launch", h.value, st)
            e.record(side)
            torch.cuda.synchronize()
            us = 1e3 * s.elapsed_time(e) / reps
        finally:
            L.call("tg_graph_destroy", h.value)
        per_step = counter.counts[(name, mode)] / steps
        out[f"{name}/{EP_NAMES[mode]}"] = {
            "avg_us": us, "bytes_per_launch": nb, "gbps": nb / (us * 1e-6) / 1e9,
            "frac_hbm": nb / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS, "launches_per_step": per_step,
            "ms_per_step": 1e-3 * us * per_step}
    return out


def kernel_probe(precision: str, reps: int = 20) -> dict:
    """For hardware-counter runs: two SIMP iterations of the 1M-element
    cantilever, a 256 MiB device copy of known size (what a counter run
    scales its byte counts with), then reps launches each, on the live
    operands, of the level-0 smoother steps CHEB_R, CHEB_FIRST, RESID_UPDATE,
    RESID_FIRST and CHEB_LAST, the split-BF16 PCG product (DOT) and the plain product
    (STORE), the level-1 Galerkin operator's two smoothing steps, and the
    CHEB_R step of the stencil levels 2 and 3, in that order."""
    import torch

    from topogpu_amd.hip import lib as L
    from topogpu_amd.hip.gmg import EP_NAMES, kernel_key, op_bytes

    eng = make_engine("c2", precision)
    for _ in range(2):
        eng.step()
    copy_bytes = 256 << 20
    src = torch.ones(copy_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    L.call("tg_copy32", src.numel(), L.ptr(dst), L.ptr(src), st)   # k_copy32: reads and writes copy_bytes
    g = eng.fea.gmg
    order = []

    def launches(lv, op, u, y, ep):
        Lv = g.levels[lv]
        for _ in range(reps):
            L.call("tg_apply32", op, L.ptr(u), L.ptr(y), ep, st)
        order.append({"kernel": f"{kernel_key(lv, op.kind)}/{EP_NAMES[ep.mode]}", "grid_nodes": Lv.nodes,
                      "algorithmic_bytes": op_bytes(op.kind, Lv.n, Lv.n_elem, ep.mode)})

    for lv in (0, 1, 2, 3):
        Lv = g.levels[lv]
        u, r, x, d = Lv.da.clone(), Lv.r.clone(), Lv.x.clone(), torch.zeros_like(Lv.da)
        rhs, y = Lv.r.clone(), torch.zeros_like(Lv.da)
        dot = torch.zeros(1, dtype=torch.float64, device=y.device)
        op = g.op(lv)
        cheb_r = L.ep(L.TG_EP_CHEB_R, dinv=Lv.dinv, d=d, x=x, r=r, a1=0.1, a2=0.1)
        if lv == 0:
            launches(0, op, u, None, cheb_r)
            launches(0, op, rhs, None, L.ep(L.TG_EP_CHEB_FIRST, rhs=rhs, dinv=Lv.dinv, d=d, x=x, r=r, a1=0.1,
                                             a2=0.1, a3=0.5))
            launches(0, op, u, None, L.ep(L.TG_EP_RESID_UPDATE, r=r))
            launches(0, op, rhs, None, L.ep(L.TG_EP_RESID_FIRST, rhs=rhs, dinv=Lv.dinv, x=x, r=r, a3=0.5))
            launches(0, op, u, None, L.ep(L.TG_EP_CHEB_LAST, dinv=Lv.dinv, d=d, x=x, r=r, a1=0.1, a2=0.1))
            launches(0, eng.fea.fine_op(), u, y, L.ep(L.TG_EP_DOT, dot=dot))
            launches(0, op, u, y, L.ep(L.TG_EP_STORE))
        elif lv == 1:
            launches(1, op, rhs, None, L.ep(L.TG_EP_RESID_FIRST, rhs=rhs, dinv=Lv.dinv, x=x, r=r, a3=0.5))
            launches(1, op, u, None, L.ep(L.TG_EP_CHEB_LAST, dinv=Lv.dinv, d=d, x=x, r=r, a1=0.1, a2=0.1))
        else:
            launches(lv, op, u, None, cheb_r)
    torch.cuda.synchronize()
    return {"copy_bytes": copy_bytes, "copy_kernel": "k_copy32", "launches_each": reps, "order": order}


def matvec_bench(reps: int = 200, warmup: int = 5):
    """E4 protocol of the reference (run_experiments_e1_e10.py:604-623):
    5 warm-up products then 200 timed, here per precision, with the
    relative error against the FP64 GPU operator on the same inputs."""
    import torch

    from topogpu_amd import fem
    from topogpu_amd.hip import lib as L

    dims = C1
    dev = torch.device("cuda", 0)
    L.init_device(fem.unit_stiffness(), 0)
    n_nodes = int(np.prod([d + 1 for d in dims]))
    ndof, ne = 3 * n_nodes, int(np.prod(dims))
    mask = np.full(n_nodes, 7, np.uint8)
    mask.reshape(dims[0] + 1, -1)[0] = 0                         # x=0 face clamped
    rho = np.random.default_rng(0).uniform(1e-3, 1.0, ne)
    E = rho ** 3
    u = np.random.default_rng(1).standard_normal(ndof)
    free = np.repeat(mask, 3).reshape(-1, 3) >> np.arange(3) & 1
    u = u * free.ravel()
    dmask = torch.from_numpy(mask).to(dev)
    E64, u64 = torch.from_numpy(E).to(dev), torch.from_numpy(u).to(dev)
    E32, u32 = E64.float(), u64.float()
    y64 = torch.zeros(ndof, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L.call("tg_apply64", L.grid(dims), L.ptr(dmask), L.ptr(E64), L.ptr(u64), None, L.ptr(y64), None, None, st)
    ref = y64.cpu().numpy()
    out = {"grid": "x".join(map(str, dims)), "n_elem": ne, "ndof": ndof,
           "bytes_per_matvec": 8 * ndof + 4 * ne, "flops_per_matvec": 1152 * ne}
    kinds = {"fp32": L.TG_OP_HEX_F32, "bf16": L.TG_OP_HEX_BF16, "bf16x3": L.TG_OP_HEX_BF16X3}
    y = torch.zeros(ndof, dtype=torch.float32, device=dev)
    for name, kind in kinds.items():
        op = L.TgOp32(kind, L.grid(dims), L.ptr(dmask), L.ptr(E32), None)
        ep = L.ep(L.TG_EP_STORE)
        for _ in range(warmup):
            L.call("tg_apply32", op, L.ptr(u32), L.ptr(y), ep, st)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            L.call("tg_apply32", op, L.ptr(u32), L.ptr(y), ep, st)
        e.record()
        torch.cuda.synchronize()
        us = 1e3 * s.elapsed_time(e) / reps
        got = y.cpu().numpy().astype(np.float64)
        out[name] = {"us": us, "gbps": out["bytes_per_matvec"] / (us * 1e-6) / 1e9,
                     "tflops": out["flops_per_matvec"] / (us * 1e-6) / 1e12,
                     "rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                     "rel_inf_vs_fp64": float(np.abs(got - ref).max() / np.abs(ref).max())}
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(20):
        L.call("tg_apply64", L.grid(dims), L.ptr(dmask), L.ptr(E64), L.ptr(u64), None, L.ptr(y64), None, None, st)
    e.record()
    torch.cuda.synchronize()
    us = 1e3 * s.elapsed_time(e) / 20
    out["fp64"] = {"us": us, "gbps": (16 * ndof + 8 * ne) / (us * 1e-6) / 1e9}
    return out


def cpu_c0_baseline(min_seconds: float = 10.0) -> dict:
    """The reference CPU algorithm (run_simp: filter, Heaviside, direct
    solve, filtered-volume OC; the oracle's FP64 restatement, pinned to the
    reference's own trajectories) on the 24x12x6 cantilever, 3 OC
    iterations, repeated until min_seconds of CPU work, with BLAS limited to
    one thread (SuperLU is single-threaded): iterations per second. The GPU
    runs the same case through SIMPSolver(backend="hip") beside it."""
    import torch

    import oracle
    import topogpu_amd as tg
    from threadpoolctl import threadpool_limits
    from topogpu_amd.bc import generate_bc

    prob = tg.unit_cantilever((24, 12, 6), volfrac=0.30, rmin=1.5)
    bc = generate_bc(prob.spec)
    kw = dict(rmin=1.5, move=0.2, max_iter=3, min_iter=2, tol=1e-3)
    with threadpool_limits(limits=1):
        oracle.run_simp(24, 12, 6, 0.30, bc.free_dofs, bc.F, **kw)          # warm
        runs = its = 0
        t0 = time.perf_counter()
        while True:
            r = oracle.run_simp(24, 12, 6, 0.30, bc.free_dofs, bc.F, **kw)
            runs += 1
            its += int(r["n_iter"])
            dt = time.perf_counter() - t0
            if dt >= min_seconds and runs >= 2:
                break
    cpu = its / dt
    solver = tg.SIMPSolver(backend="hip", tol=1e-5, max_iter=3, move=0.2)
    solver.solve(prob)                                                    # warm (library load, first launches)
    torch.cuda.synchronize()
    g_runs, g_its = 5, 0
    t0 = time.perf_counter()
    for _ in range(g_runs):
        g_its += int(solver.solve(prob).summary["n_iter"])
    torch.cuda.synchronize()
    gpu = g_its / (time.perf_counter() - t0)
    return {"value": cpu, "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": (f"24x12x6 cantilever (x=0 clamped, unit -z load at node (24, 6, 0), volfrac 0.3, rmin 1.5, "
                       f"move 0.2), 3 OC iterations per run_simp call with its final re-evaluation; {runs} calls, "
                       f"{its} iterations in {dt:.1f} s; FP64 oracle restatement of the reference CPU path "
                       f"(filter, Heaviside beta=1, scipy spsolve, filtered-volume OC), BLAS threads 1"),
            "gpu_same_sample": {"value": gpu, "unit": "iter/s",
                                "sample": f"SIMPSolver(backend='hip', tol 1e-5), same case, {g_runs} whole runs"}}


def cpu_c2_model(fine_ops_per_iter: float, sample_matvecs: int = 2) -> dict:
    """A model, not a timed run: the oracle's FP64 matrix-free operator
    (gather, GEMM with KE, bincount scatter -- MatrixFreeKff restated) timed
    at 160x80x80, times the fine-operator applications per SIMP iteration
    counted on the GPU run, plus one host filter / sensitivity / OC pass at
    the same size (coarse levels and the coarsest solve are not charged). A
    direct solve of 3.15M DOF does not fit on the host."""
    import oracle
    from oracle import simp as osimp

    try:
        from threadpoolctl import threadpool_info
        blas_threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    except Exception:  # pragma: no cover
        blas_threads = 1
    dims = C2["dims"]
    ne = int(np.prod(dims))
    ndof = 3 * int(np.prod([d + 1 for d in dims]))
    edof = oracle.edof_table_3d(*dims)
    rng = np.random.default_rng(0)
    E = rng.uniform(1e-3, 1.0, ne) ** 3
    fm = np.ones(ndof, bool)
    fm[: 3 * 81 * 81] = False
    u = rng.standard_normal(ndof)
    osimp.matfree_matvec(u, E, edof, fm)                        # warm
    t0 = time.perf_counter()
    for _ in range(sample_matvecs):
        osimp.matfree_matvec(u, E, edof, fm)
    t_mv = (time.perf_counter() - t0) / sample_matvecs
    H = oracle.build_density_filter(*dims, C2["rmin"])
    rho = np.full(ne, C2["volfrac"])
    t0 = time.perf_counter()
    rf = H @ rho
    rp = oracle.heaviside(rf, 1.0)
    dh = oracle.heaviside_deriv(rf, 1.0)
    ce = oracle.element_compliance(u, edof)
    dc = H.T @ (-3.0 * rp ** 2 * ce * dh)
    dv = H.T @ dh
    oracle.oc_update_filtered(rho, dc, dv, C2["volfrac"], C2["move"], H, 1.0, 0.5, True)
    t_host = time.perf_counter() - t0
    per_iter = fine_ops_per_iter * t_mv + t_host
    return {"kind": "model", "value": 1.0 / per_iter, "unit": "iter/s", "cores": os.cpu_count(),
            "threads": {"blas": int(blas_threads), "gather_scatter": 1},
            "sample": (f"160x80x80: {sample_matvecs} FP64 matrix-free matvecs of the oracle ({t_mv:.3f} s each) "
                       f"x {fine_ops_per_iter:.1f} fine-operator applications per SIMP iteration (GPU run) + one "
                       f"filter/sensitivity/OC pass ({t_host:.2f} s); coarse levels not charged"),
            "t_matvec_s": t_mv, "t_simp_host_s": t_host}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48, help="timed SIMP iterations (default: with the warm-up, "
                                                           "the 50 OC iterations of the workload)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="bf16")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-matvec", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the 4-level, degree-3 solver run")
    ap.add_argument("--no-multi-load", action="store_true", help="skip the 8-load bracket block run")
    ap.add_argument("--levels", default=None, help="override the hierarchy depth of the headline run (int or auto)")
    ap.add_argument("--matvec-only", action="store_true", help="only the 128x64x64 matvec microbenchmark")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="two warm-up SIMP iterations, then a 256 MiB device copy and 20 launches each of the "
                         "level-0, level-1 and stencil-level kernels on the live operands (for hardware-counter "
                         "runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this path has no multi-GPU mode; run with --gpus 1", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 2
    if args.matvec_only:
        print(json.dumps({"matvec": matvec_bench()}), flush=True)
        return 0
    if args.kernel_probe:
        print(json.dumps({"kernel_probe": kernel_probe(args.precision)}), flush=True)
        return 0

    levels = None if args.levels is None else (args.levels if args.levels == "auto" else int(args.levels))
    bf, launches, kernels, sizes = simp_bench(args.precision, args.steps, args.warmup, n_levels=levels,
                                              saturated=True)
    _log(f"{args.precision}: {bf['iter_per_s']:.3f} iter/s ({bf['n_levels']} levels)")
    fp = None
    if not args.no_fp32:
        fp, _, _, _ = simp_bench("fp32", args.steps, args.warmup, n_levels=levels)
        _log(f"fp32: {fp['iter_per_s']:.3f} iter/s")
    variants = {}
    if not args.no_variants:
        # the solver the workload description names: 4 levels (coarsest
        # 7,260 free DOFs on 21 x-planes, factored each iteration by block
        # cyclic reduction), Chebyshev degree 3
        st, _, _, _ = simp_bench(args.precision, args.steps, args.warmup, n_levels=4, fine_degree=3,
                                 coarse_degree=3)
        variants[f"{args.precision}_4_levels_degree_3"] = st
        _log(f"{args.precision} 4 levels, degree 3: {st['iter_per_s']:.3f} iter/s")
    multi = None
    if not args.no_multi_load:
        # 30 OC iterations of the 8-load bracket (2 warm-up + 28 timed), the
        # block of 8 right-hand sides against its first load case alone
        blk, _, _, sz4 = simp_bench(args.precision, 28, 2, case="c4")
        one, _, _, _ = simp_bench(args.precision, 28, 2, case="c4", loads=[0])
        multi = {"workload": "multi-load bracket 192x64x64 hex8 (786,432 elem), y=0 clamped, 8 unit point loads "
                             "on the y=64 face (default_rng(3)), equal weights, 8-RHS block PCG-GMG, tol 1e-5, "
                             "OC iterations 3-30 of 30 timed",
                 "block_8_loads": blk, "single_load_0": one, "ndof": sz4["ndof"],
                 "load_case_solves_per_s": 8 * blk["iter_per_s"],
                 "block_cost_in_single_loads": one["iter_per_s"] / blk["iter_per_s"]}
        _log(f"multi-load: block {blk['iter_per_s']:.3f} iter/s, single load {one['iter_per_s']:.3f} iter/s")
    mv = None if args.no_matvec else matvec_bench()
    # dominant kernel: the accounted (level, kernel, epilogue) launch with
    # the most device time per SIMP step, over every level and transfer
    ranked = sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_step"])
    dom_name, dom = ranked[0]
    full = C2["dims"] == (160, 80, 80)
    traffic, traffic_src = profiled_value(dom_name, "traffic_bytes_per_launch") if full else (None, None)
    mfma_busy, mfma_src = profiled_value(dom_name, "mfma_busy_fraction") if full else (None, None)
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": dom["gbps"], "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": dom["gbps"] / HBM_PEAK_GBPS, "traffic": traffic,
                "traffic_source": traffic_src, "mfma_busy_fraction": mfma_busy, "mfma_source": mfma_src,
                "avg_us": dom["avg_us"], "launches_per_step": dom["launches_per_step"],
                "algorithmic_bytes_per_launch": dom["bytes_per_launch"],
                "device_ms_per_step": dom["ms_per_step"]}
    coarse = [(k, v) for k, v in ranked if not k.startswith("L0 ") and not k.startswith("L1 ")]
    if coarse:
        ck, cv = coarse[0]
        roofline["coarse_levels"] = {"kernel": ck, "achieved": cv["gbps"], "frac": cv["frac_hbm"],
                                     "avg_us": cv["avg_us"], "launches_per_step": cv["launches_per_step"],
                                     "device_ms_per_step": cv["ms_per_step"],
                                     "all_coarse_operator_ms_per_step": sum(v["ms_per_step"] for _, v in coarse)}
    # beside it, the level-0 operator launch with the most device time (the
    # MFMA kernel the precision choice is about) and its counter figures
    fine = [(k, v) for k, v in ranked if k.startswith("L0 hex")]
    if fine:
        fk, fv = fine[0]
        ft, fts = profiled_value(fk, "traffic_bytes_per_launch") if full else (None, None)
        fm, fms = profiled_value(fk, "mfma_busy_fraction") if full else (None, None)
        roofline["level0"] = {"kernel": fk, "achieved": fv["gbps"], "frac": fv["frac_hbm"], "traffic": ft,
                              "traffic_source": fts, "mfma_busy_fraction": fm, "mfma_source": fms,
                              "avg_us": fv["avg_us"], "launches_per_step": fv["launches_per_step"],
                              "algorithmic_bytes_per_launch": fv["bytes_per_launch"],
                              "device_ms_per_step": fv["ms_per_step"],
                              "all_level0_operator_ms_per_step": sum(v["ms_per_step"] for _, v in fine)}
    roofline["accounted_ms_per_step"] = sum(v["ms_per_step"] for v in kernels.values())
    fine_ops = sum(v for k, v in launches.items() if k.startswith("L0 hex"))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_c0_baseline()
        cpu["model_1M"] = cpu_c2_model(fine_ops)
    w0, w1 = args.warmup + 1, args.warmup + args.steps
    line = {
        "metric": METRIC, "value": bf["iter_per_s"], "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": bf["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16" if args.precision.startswith("bf16") else "fp32",
        "data": "synthetic",
        "config": {"workload": "SIMP cantilever 160x80x80 hex8 (1,024,000 elem), x=0 clamped, line load on the "
                               f"x=160 z=0 edge, Galerkin GMG-PCG (Chebyshev degree {degree_schedule(C2)}), tol 1e-5; "
                               f"OC iterations {w0}-{w1} of a run of {w1} from rho=volfrac timed",
                   "precision": args.precision, **{k: v for k, v in C2.items() if k != "dims"},
                   "n_levels_used": bf["n_levels"], "coarsest_dofs": bf["coarsest_dofs"],
                   "grid": "160x80x80", "ndof": sizes["ndof"], "n_elem": sizes["n_elem"]},
        "solve": {k: bf[k] for k in ("pcg_iters_mean", "pcg_iters_max", "ir_restarts_mean", "max_rel_residual",
                                     "all_converged", "compliance_last")},
        "fp32_plain": fp, "speedup_vs_fp32": (bf["iter_per_s"] / fp["iter_per_s"]) if fp else None,
        "variants": variants, "multi_load": multi, "launches_per_step": launches, "matvec": mv,
        "kernels": dict(ranked[:12]), "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
