#!/usr/bin/env python3
"""AMG-PCG time-to-1e-8 on one MI355X: FP32-cycle vs FP64, V-cycle HBM GB/s.

Workload (default): 3-D 27-point Laplacian on a 256^3 grid (16,777,216 rows,
449,455,096 nonzeros), Ruge-Stuben setup (classical strength 0.25, PMIS,
extended interpolation with filter 0.3), weighted Jacobi w = 2/3, b ~ N(0,1)
(seed 2).  One step = one complete solve from x = 0 to a true FP64 relative
residual ||b - Ax|| / ||b|| <= 1e-8 with flexible PCG in FP64 and the
V-cycle in FP32 (the headline `value`, seconds, lower is better); the same
solve with the V-cycle in FP64 is timed beside it.  `--config c2` runs the
128^3 grid with b = A x_true, x_true ~ U(-1,1) (seed 1).

Also reported: live HIP-event timing of the dominant kernel (roofline); the
per-level timing of the setup's two sparse products (galerkin); and, outside
every timed GPU region, the oracle's whole solve of the same system on the
same hierarchy on the host's cores (cpu_baseline, with the one-core figure
beside it) and the error of the GPU solutions against the oracle's FP64 solve
to 1e-12 (oracle_fp64_solve).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMG-PCG time-to-1e-8 residual (s), FP32-cycle vs FP64; V-cycle HBM GB/s"
PEAK_HBM_GBS = 8000.0
CONFIGS = {
    "c2": dict(m=128, rhs="xtrue_uniform", seed=1),
    "c3": dict(m=256, rhs="normal", seed=2),
    "c4": dict(m=2048, k=32, seed=100),
    "c5": dict(m=96, coef_seed=3, xstar_seed=4, fd_seed=5, rhs_seed=6),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def vcycle_bytes(ml, sv):
    """Algorithmic HBM bytes of one V(1,1) cycle from x = 0 with the value
    width sv (8 FP64, 4 FP32), int32 indices, per level:
      pre-smooth from zero  x = w b / d           3 sv n
      residual  r = b - A x                        nnz (sv+4) + 4(n+1) + 3 sv n
      restrict  b_c = R r                          nnzP (sv+4) + 4(nc+1) + sv nf + sv nc
      interpolate x += P e                         nnzP (sv+4) + 4(nf+1) + sv nc + 2 sv nf
      post-smooth Jacobi                           nnz (sv+4) + 4(n+1) + 3 sv n
    plus the coarse FP64 LU solve 8 nc^2 + 16 nc.  A level held in the
    stencil-diagonal form streams nd*n values and a 4-byte mask per row
    instead of nnz (sv+4) + 4(n+1)."""
    n, nnzA, nnzP = ml.level_sizes()
    total = 0.0
    L = len(n)
    for l in range(L - 1):
        nf, nc = int(n[l]), int(n[l + 1])
        nd = ml.level(l)[0].diagonals()
        a = nd * nf * sv + 4 * nf if nd else nnzA[l] * (sv + 4) + 4 * (nf + 1)
        p = nnzP[l] * (sv + 4)
        total += 3 * sv * nf
        total += a + 3 * sv * nf
        total += p + 4 * (nc + 1) + sv * nf + sv * nc
        total += p + 4 * (nf + 1) + sv * nc + 2 * sv * nf
        total += a + 3 * sv * nf
    nc = int(n[-1])
    total += 8 * nc * nc + 16 * nc
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--m", type=int, default=0, help="override the grid edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-cycle comparison solves")
    ap.add_argument("--no-galerkin", action="store_true", help="skip the per-level SpGEMM timing")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (default: OMP_NUM_THREADS or the CPUs this "
                         "process may use, at most 16)")
    ap.add_argument("--no-graph", action="store_true",
                    help="block solves (c4): launch every iteration eagerly instead of replaying captured HIP graphs")
    ap.add_argument("--graphs", action="store_true",
                    help="single-vector solves: replay iterations from captured HIP graphs (default: eager launches, "
                         "which measured faster)")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with one axpy over 2^27 doubles (2 GiB read, 1 GiB written), the reference "
                         "dispatch for scaling hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import raptor_amd as ra
    from raptor_amd import _lib, synth

    cfg = dict(CONFIGS[args.config])
    m = args.m or cfg["m"]
    ra.Context.graphs_default = bool(args.graphs)
    if args.config == "c4":
        return bench_batch(args, m, cfg)
    if args.config == "c5":
        return bench_adjoint(args, m, cfg)
    ctx = ra.Context(0)
    t = time.time()
    rp, col, val = ra.gallery_laplace27(m)
    n = rp.size - 1
    log(f"27-point Laplacian {m}^3: n={n} nnz={col.size} generated in {time.time() - t:.1f}s")

    t = time.time()
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.PMIS, ra.Extended, ra.Classical, ra.Jacobi)
    ml.relax_weight = 2.0 / 3.0
    ml.setup(rp, col, val)
    setup_s = time.time() - t
    tphase = ml.setup_times()
    nl, nnzA, nnzP = ml.level_sizes()
    log(f"setup {setup_s:.1f}s (strength {tphase[0]:.1f} split {tphase[1]:.1f} interp {tphase[2]:.1f} "
        f"galerkin {tphase[3]:.1f} other {tphase[4]:.1f}); levels {list(map(int, nl))}")
    ml.enable_fp32()
    caps = ml.row_caps()
    log(f"row-block caps [A, P, R] x [fp64, fp32] by level: {caps.tolist()}")
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)

    if cfg["rhs"] == "normal":
        b_h = synth.normal(cfg["seed"], n)
    else:
        xt = synth.uniform_range(cfg["seed"], n, -1.0, 1.0)
        xd = ra.ParVector.from_numpy(ctx, xt)
        bd = ra.ParVector(ctx, n)
        A.mult(xd, bd)
        b_h = bd.numpy()
        del xd, bd
    b = ra.ParVector.from_numpy(ctx, b_h)
    x = ra.ParVector(ctx, n)

    def solve(prec):
        x.set_const_value(0.0)
        hist, it, rel = ra.FPCG(A, ml, x, b, 1e-8, 500, prec)
        return it, rel

    for _ in range(args.warmup):
        solve(ra.PREC_FP32)
        if not args.no_fp64:
            solve(ra.PREC_FP64)
    log("warmup done")

    L = _lib.lib()
    fine_nnz = int(nnzA[0])
    check = _lib.check
    # ---- timed: FP32-cycle solves (headline; eager launches, or iterations
    # replayed from captured HIP graphs with --graphs)
    ctx.sync()
    t0 = time.perf_counter()
    its32 = []
    for _ in range(args.steps):
        it, rel32 = solve(ra.PREC_FP32)
        its32.append(it)
    ctx.sync()
    t32 = (time.perf_counter() - t0) / args.steps
    # ---- the same solves again with the fine-level kernels bracketed by HIP
    # events on the library's stream (eager launches), for the roofline
    check(L.ramd_prof_enable(ctx.h, fine_nnz))
    for _ in range(args.steps):
        solve(ra.PREC_FP32)
    prof = {}
    import ctypes as C
    names = {0: "spmv", 1: "append", 2: "residual", 3: "jacobi", 4: "add_into"}
    for op in range(5):
        for prec in (ra.PREC_FP64, ra.PREC_FP32):
            nl_, nd_, ms, by = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
            check(L.ramd_prof_read(ctx.h, op, prec, C.byref(nl_), C.byref(nd_), C.byref(ms), C.byref(by)))
            if nl_.value:
                kern = "dia_kernel" if nd_.value == nl_.value else ("stream_kernel" if nd_.value == 0 else "mixed")
                prof[f"{names[op]}_{'fp32' if prec else 'fp64'}"] = dict(
                    kernel=kern, launches=nl_.value, ms=ms.value, bytes=by.value,
                    gbs=round(by.value / (ms.value * 1e-3) / 1e9, 1))
    check(L.ramd_prof_enable(ctx.h, -1))
    log(f"FP32-cycle solve: {t32:.4f}s, iterations {its32}, relres {rel32:.3e}")
    x32_h = x.numpy()

    # ---- error of the FP32-cycle solution against an FP64 reference solve
    # (FP64 cycles, whose V-cycle is bit-identical to the oracle's, to 1e-12)
    x.set_const_value(0.0)
    ra.FPCG(A, ml, x, b, 1e-12, 500, ra.PREC_FP64)
    xref = x.numpy()
    err32 = float(np.linalg.norm(x32_h - xref) / np.linalg.norm(xref))
    log(f"FP32-cycle solution vs FP64 reference (1e-12): relative error {err32:.3e}")

    # ---- FP64-cycle solves (same protocol, same outer operator)
    t64 = None
    its64 = []
    rel64 = None
    x64_h = None
    if not args.no_fp64:
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            it, rel64 = solve(ra.PREC_FP64)
            its64.append(it)
        ctx.sync()
        t64 = (time.perf_counter() - t0) / args.steps
        x64_h = x.numpy()
        log(f"FP64-cycle solve: {t64:.4f}s, iterations {its64}, relres {rel64:.3e}")

    # ---- the plain FP64 form: FP64 cycles and the outer operator on A's
    # FP64 values as the reference stores them (identical results; the
    # matched comparison above gives both forms the exact FP32 value copy)
    t64p = None
    plain_same = None
    plain_its, plain_rel = [], None
    if not args.no_fp64 and A.diagonals():
        A.set_fp32_outer(False)
        solve(ra.PREC_FP64)  # the outer operator's kernel choice for FP64 values
        ctx.sync()
        t0 = time.perf_counter()
        plain_same = True
        for _ in range(args.steps):
            it, relp = solve(ra.PREC_FP64)
            plain_its.append(it)
            plain_rel = relp
            plain_same = plain_same and it == its64[-1] and relp == rel64
        ctx.sync()
        t64p = (time.perf_counter() - t0) / args.steps
        plain_same = bool(plain_same and np.array_equal(x.numpy(), x64_h))
        A.set_fp32_outer(True)
        log(f"plain FP64 form (FP64 outer values): {t64p:.4f}s, iterations {plain_its}, relres {plain_rel:.3e}, "
            f"bit-identical to the matched FP64 solve: {plain_same}")

    # ---- V-cycle throughput (FP32 and FP64), 10 cycles each
    vc = {}
    zb = ra.ParVector.from_numpy(ctx, b_h)
    x32 = ra.ParVector(ctx, n, np.float32)
    b32 = ra.ParVector.from_numpy(ctx, b_h.astype(np.float32))
    for prec_name, sv in (("fp32", 4), ("fp64", 8)):
        cyc = 10
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(cyc):
            if prec_name == "fp64":
                ml.cycle(x, zb, x_is_zero=True)
            else:
                check(L.ramd_vcycle_fp32(ml.h, x32.ptr, b32.ptr))
        ctx.sync()
        dt = (time.perf_counter() - t0) / cyc
        by = vcycle_bytes(ml, sv)
        vc[prec_name] = dict(bytes=by, s=round(dt, 6), gbs=round(by / dt / 1e9, 1))
        log(f"V-cycle {prec_name}: {dt * 1e3:.3f} ms, {by / dt / 1e9:.0f} GB/s algorithmic")

    # ---- the reference's PCG rules (krylov/par_cg.cpp:121-263, FP64 cycle) on
    # the same system, once, outside the timed region: its iteration count
    xr = ra.ParVector(ctx, n)
    ctx.sync()
    t0 = time.perf_counter()
    _, it_ref = ra.PCG(A, ml, xr, b, tol=1e-8)
    ctx.sync()
    t_ref = time.perf_counter() - t0
    rr = ra.ParVector(ctx, n)
    A.residual(xr, b, rr)
    rel_ref = float(np.linalg.norm(rr.numpy()) / np.linalg.norm(b_h))
    pcg_ref = dict(iterations=it_ref, s=round(t_ref, 5), true_relres=rel_ref)
    log(f"PCG with the reference's rules (FP64 cycle, tol 1e-8): {it_ref} iterations, {t_ref:.4f}s, "
        f"true relres {rel_ref:.3e}")
    del xr, rr

    # ---- roofline: dominant kernel (most total event time)
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    per_launch_bytes = dom["bytes"] / dom["launches"]
    per_launch_s = dom["ms"] / 1e3 / dom["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    # the outer product runs on the caller's matrix A (FP32 values, FP64
    # vectors: pair 2); the cycle's operations on level 0 (pair 1)
    outer_slot = dom_name.endswith("fp64")
    dom_op = A if outer_slot else ml.level(0)[0]
    staging = dom_op.dia_staging(2 if outer_slot else 1) if dom_op.diagonals() else None
    kname = kernel_template(dom_name, dom["kernel"], ml.level(0)[0].diagonals(), staging)
    traffic = measured_traffic(m, kname)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic, kernel=f"{dom['kernel']}:{dom_name}",
                    kernel_symbol=kname, staging=staging, bytes_per_launch=per_launch_bytes,
                    us_per_launch=round(per_launch_s * 1e6, 2))
    log(f"dominant kernel {dom_name}: {achieved:.0f} GB/s ({achieved / PEAK_HBM_GBS:.1%} of 8 TB/s)")

    # ---- Galerkin product per level (the setup's two SpGEMMs, timed alone)
    galerkin = None
    if not args.no_galerkin:
        galerkin = galerkin_report(ctx, ml)

    # ---- CPU baseline and FP64-oracle error: the oracle (plain C, every
    # host core) solves the same system on the same hierarchy
    cpu = None
    oracle = None
    if not args.no_cpu_baseline:
        try:
            cpu, oracle = cpu_baseline(ml, rp, col, val, b_h, x32_h, x64_h)
        except Exception as e:  # the checker is optional on machines without it
            log(f"cpu baseline skipped: {e}")

    out = {
        "metric": METRIC,
        "value": round(t32, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-cycle/f64-outer",
        "data": "synthetic",
        "config": dict(workload=f"27pt-laplacian-{m}^3 RS(PMIS, extended, theta=0.25) Jacobi(2/3) FPCG to 1e-8",
                       n=n, nnz=int(col.size), rhs=cfg["rhs"], seed=cfg["seed"], levels=list(map(int, nl))),
        "fp64_cycle_s": round(t64, 5) if t64 else None,
        # both forms run the same outer FP64 operator (FP32-stored values
        # where they are exact) and the same vector updates and dots
        "fp32_speedup_vs_fp64": round(t64 / t32, 3) if t64 else None,
        # against the plain FP64 form (its outer operator reads the FP64
        # values; same iterates)
        "fp64_plain_s": round(t64p, 5) if t64p else None,
        "fp32_speedup_vs_plain_fp64": round(t64p / t32, 3) if t64p else None,
        "fp64_plain_same_iterates": plain_same,
        "iterations_fp64_plain": plain_its or None,
        "relres_fp64_plain": plain_rel,
        "outer_operator": outer_operator(A),
        "launch": "graphs" if args.graphs else "eager",
        "iterations_fp32": its32,
        "iterations_fp64": its64,
        "relres_fp32": rel32,
        "relres_fp64": rel64,
        "error_fp32_vs_fp64_reference": err32,
        "oracle_fp64_solve": oracle,
        "pcg_reference_rule": pcg_ref,
        "setup_s": round(setup_s, 2),
        "vcycle": vc,
        "galerkin": galerkin,
        "kernels": prof,
        "row_caps": caps.tolist(),
        # stencil-diagonal kernel per precision pair: level 0 FP64 and FP32
        # cycles, the outer operator (mode, planes, L, P)
        "dia_staging": dict(level0_fp64=ml.level(0)[0].dia_staging(0), level0_fp32=ml.level(0)[0].dia_staging(1),
                            outer=A.dia_staging(2)) if A.diagonals() else None,
        "coarse_tail": coarse_tail(ml),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.calibrate:
        calibrate_copy(ctx)
    print(json.dumps(out), flush=True)
    ctx.close()


def kernel_template(slot, kind, nd, staging=None):
    """Symbol of the kernel behind a timing slot of the FP32-form solve: the
    FP64-vector slots there read the lossless FP32 values.  staging: the
    stencil-diagonal variant (mode, planes, L, P) of the operator the slot
    runs on (gather, LDS x window, or plane-march over the box grid)."""
    op = {"spmv": 0, "append": 1, "residual": 2, "jacobi": 3, "add_into": 4}[slot.rsplit("_", 1)[0]]
    vec = "float" if slot.endswith("fp32") else "double"
    if kind == "dia_kernel":
        mode = staging[0] if staging else 0
        if mode == 2:
            L = staging[2]
            tx = 64 if L % 64 == 0 else 32 if L % 32 == 0 else 16 if L % 16 == 0 else 64
            return f"ramd::dia_box_kernel<{op}, float, {vec}, {nd}, {tx}>"
        name = "dia_lds_kernel" if mode == 1 else "dia_kernel"
        return f"ramd::{name}<{op}, float, {vec}, {nd}>"
    return f"ramd::stream_kernel<{op}, {vec}, {vec}>"


def measured_traffic(m, kname):
    """HBM bytes per launch of kname from the committed counter summary of
    the same workload (profiles/pmc_traffic_27pt-<m>cubed_fpcg.md), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_27pt-{m}cubed_fpcg.md")
    try:
        txt = open(path).read()
        block = txt.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block).get(kname)
    except (OSError, IndexError, ValueError):
        return None


def calibrate_copy(ctx):
    """One ramd_axpy (y += a x) over 2^27 doubles: 2 GiB read, 1 GiB written;
    its FETCH_SIZE / WRITE_SIZE counters scale the other kernels' counters."""
    import raptor_amd as ra
    n = 1 << 27
    x, y = ra.ParVector(ctx, n), ra.ParVector(ctx, n)
    x.set_const_value(1.0)
    y.set_const_value(0.0)
    ctx.sync()
    y.axpy(x, 0.5)
    ctx.sync()
    log(f"calibration axpy: n={n} (reads {16 * n} B, writes {8 * n} B)")
    x.free()
    y.free()


def _levels_of(ml):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as K
    As, Ps = [], []
    for l in range(ml.num_levels):
        Al, Pl = ml.level(l)
        As.append(K.Csr(*Al.to_csr(), Al.n_cols))
        if Pl is not None:
            Ps.append(K.Csr(*Pl.to_csr(), Pl.n_cols))
    return As, Ps


def bench_batch(args, m, cfg):
    """2-D anisotropic diffusion (eps 0.001, theta pi/8) on m x m, RS(CLJP,
    modified classical, 0.25), Jacobi 2/3; k right-hand sides b_c ~ N(0,1)
    (seeds 100..100+k-1) solved together by the lockstep block PCG (graph
    replay, FP64 and FP32 cycles) against k sequential single-RHS solves.
    One step = all k systems to 1e-8 true relative residual."""
    import numpy as np

    import raptor_amd as ra
    from raptor_amd import synth
    k = cfg["k"]
    ctx = ra.Context(0)
    t = time.time()
    rp, col, val = ra.gallery_aniso2d(m, 0.001, math.pi / 8)
    n = rp.size - 1
    log(f"anisotropic 9-point {m}^2: n={n} nnz={col.size} generated in {time.time() - t:.1f}s")
    t = time.time()
    ml = ra.ParRugeStubenSolver(ctx, 0.25, ra.CLJP, ra.ModClassical, ra.Classical, ra.Jacobi)
    ml.relax_weight = 2.0 / 3.0
    ml.setup(rp, col, val)
    setup_s = time.time() - t
    nl, nnzA, nnzP = ml.level_sizes()
    log(f"setup {setup_s:.1f}s; levels {list(map(int, nl))}; nnz(A) {list(map(int, nnzA))}; "
        f"nnz(P) {list(map(int, nnzP))}")
    ml.enable_fp32()
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    Bh = np.stack([synth.normal(cfg["seed"] + c, n) for c in range(k)], axis=1)
    B = ra.ParVector.from_numpy(ctx, Bh.ravel())
    X = ra.ParVector(ctx, n * k)
    x1 = ra.ParVector(ctx, n)
    cols = [ra.ParVector.from_numpy(ctx, Bh[:, c].copy()) for c in range(k)]
    tuning = ra.block_tuning(ml, k)
    tuning32 = ra.block_tuning(ml, k, ra.PREC_FP32)
    log(f"block kernels (A, P, R) by level, lanes per row (negative: 16-byte column units): "
        f"FP64 {tuning.tolist()}; FP32 {tuning32.tolist()}")

    def block(prec):
        X.set_const_value(0.0)
        return ra.FPCG_block(A, ml, X, B, k, 1e-8, 500, prec, use_graph=not args.no_graph)

    def sequential(prec):
        its = []
        for c in range(k):
            x1.set_const_value(0.0)
            _, it, _ = ra.FPCG(A, ml, x1, cols[c], 1e-8, 500, prec)
            its.append(it)
        return its

    for _ in range(args.warmup):
        block(ra.PREC_FP32)
        block(ra.PREC_FP64)
        sequential(ra.PREC_FP32)
    res = {}
    for name, fn, prec in (("block_fp32", block, ra.PREC_FP32), ("block_fp64", block, ra.PREC_FP64),
                           ("sequential_fp32", sequential, ra.PREC_FP32),
                           ("sequential_fp64", sequential, ra.PREC_FP64)):
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn(prec)
        ctx.sync()
        dt = (time.perf_counter() - t0) / args.steps
        its = out[0].tolist() if name.startswith("block") else out
        res[name] = dict(s=round(dt, 5), rhs_per_s=round(k / dt, 2), iterations=[int(v) for v in its])
        log(f"{name}: {dt:.4f}s for {k} RHS ({k / dt:.1f} RHS/s), iterations {min(its)}..{max(its)}")
    # true residuals of the last block FP32 solve
    block(ra.PREC_FP32)
    Xh = X.numpy().reshape(n, k)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    Ah = None
    worst = 0.0
    try:
        import checkers as K
        Ah = K.Csr(rp, col, val)
        for c in range(k):
            worst = max(worst, float(np.linalg.norm(Bh[:, c] - K.spmv(Ah, Xh[:, c])) / np.linalg.norm(Bh[:, c])))
    except Exception as e:  # the checker is optional
        log(f"residual check skipped: {e}")
    out = {
        "metric": "block AMG-PCG time per right-hand side to 1e-8 (s), 32 RHS",
        "value": round(res["block_fp32"]["s"] / k, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["block_fp32"]["s"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-cycle/f64-outer",
        "data": "synthetic",
        "config": dict(workload=f"aniso9pt-{m}^2 RS(CLJP, mod-classical, theta=0.25) Jacobi(2/3) "
                                f"{k}-RHS lockstep FPCG to 1e-8", n=n, nnz=int(col.size), k=k,
                       levels=list(map(int, nl)), graphs=not args.no_graph),
        "setup_s": round(setup_s, 2),
        "solves": res,
        "block_speedup_fp32": round(res["sequential_fp32"]["s"] / res["block_fp32"]["s"], 3),
        "block_speedup_fp64": round(res["sequential_fp64"]["s"] / res["block_fp64"]["s"], 3),
        "worst_true_relres": worst,
        "block_kernels": dict(fp64=tuning.tolist(), fp32=tuning32.tolist()),
        "nnz_A": list(map(int, nnzA)), "nnz_P": list(map(int, nnzP)),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_adjoint(args, m, cfg):
    """7-point finite-volume diffusion on m^3 (k = exp(g), g ~ N(0,1)),
    smoothed aggregation (symmetric strength 0.08, MIS-2, Jacobi prolongation
    4/3), Jacobi 2/3.  L = 0.5 ||x - x*||^2; one step = forward solve +
    adjoint solve + gradient on A's pattern (FP64 cycles, rtol 1e-10).  The
    gradient is checked by central differences (h = 1e-6) on 32 random
    stored entries, each perturbed solve done by refinement with the
    unperturbed hierarchy, and lambda against the oracle's FP64 PCG."""
    import numpy as np

    import raptor_amd as ra
    from raptor_amd import synth
    ctx = ra.Context(0)
    t = time.time()
    rp, col, val = ra.gallery_fv7(m, seed=cfg["coef_seed"])
    n = rp.size - 1
    log(f"7-point FV {m}^3: n={n} nnz={col.size} generated in {time.time() - t:.1f}s")
    t = time.time()
    ml = ra.ParSmoothedAggregationSolver(ctx, 0.08)
    ml.relax_weight = 2.0 / 3.0
    ml.setup(rp, col, val)
    setup_s = time.time() - t
    nl, nnzA, nnzP = ml.level_sizes()
    log(f"setup {setup_s:.1f}s; levels {list(map(int, nl))}")
    A = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    b_h = synth.normal(cfg["rhs_seed"], n)
    xs_h = synth.normal(cfg["xstar_seed"], n)
    b, xs = ra.ParVector.from_numpy(ctx, b_h), ra.ParVector.from_numpy(ctx, xs_h)
    x = ra.ParVector(ctx, n)
    g = ra.ParVector(ctx, n)
    rtol = 1e-10
    # gradient outputs allocated once (a training loop reuses them)
    dA_buf, db_buf = ra.ParVector(ctx, col.size), ra.ParVector(ctx, n)

    def step(prec=ra.PREC_FP64):
        x.set_const_value(0.0)
        _, itf, _ = ra.FPCG(A, ml, x, b, rtol, 500, prec)
        g.copy(x)
        g.axpy(xs, -1.0)
        dA, db, ita = ra.adjoint(A, ml, x, g, rtol=rtol, cycle_prec=prec, dA=dA_buf, db=db_buf)
        return itf, ita, dA, db

    # the same step with FP32 cycles inside the FP64 flexible PCG (reported
    # beside the FP64-cycle value)
    step(ra.PREC_FP32)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itf32, ita32, dA32, _ = step(ra.PREC_FP32)
    ctx.sync()
    dt32 = (time.perf_counter() - t0) / args.steps
    gA32 = dA32.numpy()
    log(f"forward + adjoint with FP32 cycles: {dt32 * 1e3:.2f} ms (forward {itf32} iterations, adjoint {ita32})")

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itf, ita, dA, db = step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"forward + adjoint: {dt * 1e3:.2f} ms (forward {itf} iterations, adjoint {ita})")
    xh, gA, gb = x.numpy(), dA.numpy(), db.numpy()
    grad32_diff = float(np.linalg.norm(gA32 - gA) / np.linalg.norm(gA))
    log(f"dL/dA with FP32 cycles vs FP64 cycles: relative difference {grad32_diff:.2e}")
    loss = 0.5 * float(np.sum((xh - xs_h) ** 2))

    # central differences on 32 stored entries (dense-free: perturbation
    # solves delta = A'^-1 (-s h x_j e_i) by refinement with the unperturbed
    # hierarchy, then dL = (x - x*).delta + |delta|^2 / 2)
    h = 1e-6
    rng = np.random.default_rng(cfg["fd_seed"])
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ap = ra.ParCSRMatrix.from_csr(ctx, rp, col, val)
    vdev = ra.ParVector(ctx, col.size)
    r = ra.ParVector(ctx, n)
    z = ra.ParVector(ctx, n)
    d = ra.ParVector(ctx, n)
    errs = []
    t_fd = time.time()
    for kk in rng.choice(col.size, 32, replace=False):
        i, j = int(rows[kk]), int(col[kk])
        deltas = []
        for sgn in (1.0, -1.0):
            v = val.copy()
            v[kk] += sgn * h
            vdev.copy_from(v)
            Ap.set_values(vdev)
            rhs_h = np.zeros(n)
            rhs_h[i] = -sgn * h * xh[j]
            rhs = ra.ParVector.from_numpy(ctx, rhs_h)
            d.set_const_value(0.0)
            for _ in range(3):
                Ap.residual(d, rhs, r)
                z.set_const_value(0.0)
                try:
                    ra.FPCG(A, ml, z, r, 1e-12, 500, ra.PREC_FP64)
                except ra.RamdError as e:
                    if e.code != -5:
                        raise
                d.axpy(z, 1.0)
            deltas.append(d.numpy())
        dp_, dm_ = deltas
        e = xh - xs_h
        fd = (float(e @ (dp_ - dm_)) + 0.5 * float(dp_ @ dp_ - dm_ @ dm_)) / (2 * h)
        errs.append(abs(gA[kk] - fd) / max(abs(fd), 1e-300))
    log(f"finite differences on 32 entries: max relative error {max(errs):.2e} ({time.time() - t_fd:.1f}s)")

    # lambda against the oracle's FP64 PCG on the same hierarchy (CPU, one core)
    cpu = None
    lam_err = None
    if not args.no_cpu_baseline:
        try:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import checkers as K
            As, Ps = _levels_of(ml)
            H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=2.0 / 3.0)
            Ah = K.Csr(rp, col, val)
            gh = xh - xs_h
            gmg = float(gh @ H.cycle(np.zeros(n), gh))
            t = time.perf_counter()
            lam, _, itc = H.pcg(Ah, np.zeros(n), gh, tol=1e-20 * math.sqrt(gmg), max_iter=300)
            tc = time.perf_counter() - t
            lam_err = float(np.linalg.norm(lam - gb) / np.linalg.norm(lam))
            cpu = dict(value=round(tc, 2), unit="s", cores=1, kind="port",
                       sample=f"the oracle's FP64 PCG (Jacobi-smoothed V-cycles) for the adjoint system on the "
                              f"same hierarchy, {itc} iterations, one core")
            log(f"oracle adjoint solve {tc:.1f}s ({itc} iterations); lambda relative difference {lam_err:.2e}")
        except Exception as e:
            log(f"oracle comparison skipped: {e}")
    out = {
        "metric": "AMG adjoint gradient time (forward + adjoint solve + dL/dA) (s)",
        "value": round(dt, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(workload=f"fv7-lognormal-{m}^3 SA(symmetric 0.08, MIS-2, Jacobi-P 4/3) Jacobi(2/3) "
                                f"adjoint of L=0.5||x-x*||^2", n=n, nnz=int(col.size), levels=list(map(int, nl))),
        "setup_s": round(setup_s, 2),
        "iterations_forward": itf,
        "iterations_adjoint": ita,
        "loss": loss,
        "fd_max_rel_error": max(errs),
        "lambda_rel_diff_vs_oracle": lam_err,
        "fp32_cycle": dict(s=round(dt32, 5), iterations_forward=itf32, iterations_adjoint=ita32,
                           dA_rel_diff_vs_fp64_cycles=grad32_diff),
        "coarse_tail": coarse_tail(ml),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def coarse_tail(ml):
    """First level of the coarse-tail kernel each cycle precision uses (-1:
    none) and the smallest level that could start it."""
    import raptor_amd as ra
    l64, first = ml.tail_level(ra.PREC_FP64)
    l32, _ = ml.tail_level(ra.PREC_FP32)
    return dict(fp64=l64, fp32=l32, first_possible=first)


def outer_operator(A):
    """How the FPCG's outer FP64 operator reads A's values (both forms)."""
    return "stencil-diagonal, FP32-stored exact values, FP64 arithmetic" if A.diagonals() else \
        "CSR, FP64 values"


def host_cpu():
    """(model name, CPUs this process may use)"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_threads(args_threads):
    if args_threads > 0:
        return args_threads
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env if env > 0 else min(16, host_cpu()[1])


def cpu_baseline(ml, rp, col, val, b_h, x32_h, x64_h, threads=0):
    """The oracle on the GPU box's host cores, on the product's own
    hierarchy (outside every timed GPU region):

      - cpu_baseline: one whole solve of the same system to a true relative
        residual of 1e-8 by oracle/oracle_mt.c (the oracle's V-cycle,
        bit-identical to orc_cycle, row-parallel over `threads` OpenMP
        threads, inside the FP64 flexible PCG), timed; beside it the one-core
        figure: one V-cycle + one SpMV on one thread, scaled by the
        iteration count;
      - oracle_fp64_solve: the same oracle solve run to 1e-12, and the
        relative 2-norm error of the GPU's FP32-cycle and FP64-cycle
        solutions against it."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as K
    threads = cpu_threads(threads)
    model, ncpu = host_cpu()
    As, Ps = _levels_of(ml)
    H = K.Hier(As, Ps, relax_type=0, sweeps=1, omega=2.0 / 3.0)
    A = K.Csr(rp, col, val)
    n = A.n_rows
    M = K.MtHier(H, threads)
    used = M.threads
    t = time.perf_counter()
    x8, it8, rel8 = M.fpcg(A, np.zeros(n), b_h, 1e-8, 500)
    t_solve = time.perf_counter() - t
    log(f"cpu baseline: oracle FPCG on {used} threads ({model}): {t_solve:.2f}s, {it8} iterations, "
        f"true relres {rel8:.3e}")
    t = time.perf_counter()
    xo, ito, relo = M.fpcg(A, np.zeros(n), b_h, 1e-12, 500)
    t_ref = time.perf_counter() - t
    M.free()
    oracle = dict(rtol=1e-12, iterations=ito, true_relres=relo, s=round(t_ref, 2), threads=used,
                  error_fp32_cycle=float(np.linalg.norm(x32_h - xo) / np.linalg.norm(xo)),
                  error_fp64_cycle=(float(np.linalg.norm(x64_h - xo) / np.linalg.norm(xo))
                                    if x64_h is not None else None),
                  iterations_to_1e8=it8)
    log(f"oracle FP64 solve to 1e-12 ({ito} iterations, {t_ref:.1f}s): FP32-cycle solution error "
        f"{oracle['error_fp32_cycle']:.3e}, FP64-cycle {oracle['error_fp64_cycle']}")
    # one-core figure (one cycle + one SpMV, scaled)
    M1 = K.MtHier(H, 1)
    t = time.perf_counter()
    M1.cycle(np.zeros(n), b_h)
    K.spmv(A, b_h)
    per_iter = time.perf_counter() - t
    M1.free()
    one_core = per_iter * (it8 + 1)
    log(f"cpu baseline, one core: {per_iter:.2f}s per iteration -> {one_core:.1f}s per solve")
    cpu = dict(value=round(t_solve, 3), unit="s", cores=used, kind="port", cpu_model=model, host_cpus=ncpu,
               sample=f"one whole solve of the same 27-point system to a true relative residual of 1e-8 "
                      f"({it8} iterations of the oracle's FP64 V-cycle in flexible PCG, oracle/oracle_mt.c, "
                      f"{used} OpenMP threads) on the product's hierarchy",
               single_core_s=round(one_core, 2),
               single_core_sample="one FP64 V-cycle + one SpMV of the oracle on one thread, "
                                  f"{per_iter:.3f}s, scaled by {it8}+1 iterations")
    return cpu, oracle


def galerkin_report(ctx, ml):
    """Per level l: A_l P_l and P_l^T (A_l P_l) by the device SpGEMMs, each
    timed alone (host clock around the call, device synchronised), with the
    algorithmic bytes |A| + |B| + |C| (CSR: 4 (rows + 1) + 12 nnz) of the
    numeric phase plus the integer bytes (4 (rows + 1) + 4 nnz) of the
    symbolic phase, the flops 2 x (intermediate products) and the
    intermediate-product count."""
    import ctypes as C

    import numpy as np

    from raptor_amd import _lib
    L = _lib.lib()

    def rowptr(M):
        rp = np.empty(M.n_rows + 1, np.int32)
        _lib.check(L.ramd_csr_download(M.h, rp.ctypes.data_as(C.POINTER(C.c_int32)), None, None))
        return rp

    def cols(M):
        rp = np.empty(M.n_rows + 1, np.int32)
        c = np.empty(max(M.nnz, 1), np.int32)
        _lib.check(L.ramd_csr_download(M.h, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                       c.ctypes.data_as(C.POINTER(C.c_int32)), None))
        return c[:M.nnz]

    def csr_bytes(rows, nnz):
        return 4 * (rows + 1) + 12 * nnz, 4 * (rows + 1) + 4 * nnz

    out = []
    total_s = 0.0
    for l in range(ml.num_levels - 1):
        A, P = ml.level(l)
        ctx.sync()
        t0 = time.perf_counter()
        AP = A.mult(P)
        ctx.sync()
        t1 = time.perf_counter()
        Ac = AP.mult_T(P)
        ctx.sync()
        t2 = time.perf_counter()
        lp = np.diff(rowptr(P)).astype(np.int64)
        lap = np.diff(rowptr(AP)).astype(np.int64)
        prod_ap = int(lp[cols(A)].sum())
        prod_rap = int((lp * lap).sum())
        phases = []
        for name, (ra_, na_), (rb_, nb_), (rc_, nc_), prods, dt in (
                ("A*P", (A.n_rows, A.nnz), (P.n_rows, P.nnz), (AP.n_rows, AP.nnz), prod_ap, t1 - t0),
                ("P^T*(AP)", (P.n_rows, P.nnz), (AP.n_rows, AP.nnz), (Ac.n_rows, Ac.nnz), prod_rap, t2 - t1)):
            num = sum(csr_bytes(r, z)[0] for r, z in ((ra_, na_), (rb_, nb_), (rc_, nc_)))
            sym = sum(csr_bytes(r, z)[1] for r, z in ((ra_, na_), (rb_, nb_), (rc_, nc_)))
            phases.append(dict(product=name, s=round(dt, 5), bytes=num + sym, flops=2 * prods,
                               products=prods, nnz_out=int(nc_),
                               gbs=round((num + sym) / dt / 1e9, 1), gflops=round(2 * prods / dt / 1e9, 2)))
        total_s += t2 - t0
        out.append(dict(level=l, phases=phases))
        AP.free()
        Ac.free()
    log("galerkin per level: " + "; ".join(
        f"{d['level']}: " + ", ".join(f"{p['product']} {p['s'] * 1e3:.1f} ms {p['gbs']} GB/s" for p in d["phases"])
        for d in out))
    return dict(levels=out, total_s=round(total_s, 4))


if __name__ == "__main__":
    main()
