#!/usr/bin/env python3
"""Benchmark of the explicit time-stepping hot path on one MI355X.

Workload (the configuration the metric is quoted on): variable-coefficient
conservative 7-point heat on 512^3 nodes of the unit cube, kappa(x) =
1 + 0.5 sin(2 pi x) sin(2 pi y), u0 = prod sin(pi x_p) + seeded uniform noise
of amplitude 1e-3, zero Dirichlet border, tau = 0.9 h^2 / (6 kappa_max),
1000 forward Euler steps.  A "step" is one time step over all 512^3 nodes.
`value` is the FP32 + Kahan path in GLUPS (1e9 node updates per second);
the FP64 path's GLUPS and the FP32 error against it are reported beside it.

Prints ONE JSON line.  Only --gpus 1 is supported: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLUPS & %HBM peak, FP32 vs FP64 at stated error; gradient/forward time ratio"
HBM_PEAK_GBS = 8000.0
N = 512
BYTES_PER_LUP = {"f32_kahan": 20, "f32": 12, "f64": 24}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ensure_built():
    from masterthesis_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


def run_mode(ctx, capi, g, mode, dtype, u_init, k_init, tau, nsteps, warmup, slot0=0):
    """Allocate, copy the initial state, warm up, time nsteps; returns
    (ms_total, final_u DeviceArray, final_c DeviceArray or None, buffers)."""
    ne = capi.memall_count(g, 1)
    code = capi.F64 if dtype == np.float64 else capi.F32
    a = ctx.alloc(ne, dtype)
    b = ctx.alloc(ne, dtype)
    k = ctx.alloc(ne, dtype)
    for d in (a, b, k):
        d.fill(0.0)
    capi.check(ctx.L.scafes_convert(ctx.h, a.ptr, code, u_init.ptr, capi.F64, ne))
    capi.check(ctx.L.scafes_convert(ctx.h, b.ptr, code, u_init.ptr, capi.F64, ne))
    capi.check(ctx.L.scafes_convert(ctx.h, k.ptr, code, k_init.ptr, capi.F64, ne))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.alloc(ne, np.float32), ctx.alloc(ne, np.float32)
        c1.fill(0.0)
        c2.fill(0.0)
    which = capi.C.c_int(0)
    gp = capi.C.byref(g)

    def run(ns, ua, ub, ca, cb):
        capi.check(ctx.L.scafes_heat7var_run(ctx.h, gp, 1, mode, ua.ptr, ub.ptr, ca.ptr if ca else None,
                                             cb.ptr if cb else None, k.ptr, tau, ns, capi.BORDER_VALUE,
                                             0.0, -1, capi.C.byref(which)))
        return which.value

    # warm-up on scratch copies so the timed run starts from the initial state
    if warmup > 0:
        run(warmup + (warmup % 2), a, b, c1, c2)
        capi.check(ctx.L.scafes_convert(ctx.h, a.ptr, code, u_init.ptr, capi.F64, ne))
        if c1 is not None:
            c1.fill(0.0)
    ctx.sync()
    ctx.record(slot0)
    w = run(nsteps, a, b, c1, c2)
    ctx.record(slot0 + 1)
    ms = ctx.elapsed_ms(slot0, slot0 + 1)
    ctx.sync()
    res = b if w else a
    cres = (c2 if w else c1) if c1 is not None else None
    return ms, res, cres, [a, b, k, c1, c2]


def nodes(arr, n):
    return arr.reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]


def cpu_baseline(u_host, k_host, tau, n, budget_s=12.0):
    """The oracle's FP64 OpenMP step (restated reference CPU path) on the same
    512^3 input, a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    L = O.lib()
    threads = min(os.cpu_count() or 1, 16)
    L.oracle_set_threads(threads)
    g = O.make_grid((n, n, n))
    a = np.ascontiguousarray(u_host)
    b = a.copy()
    t0 = time.perf_counter()
    L.oracle_heat7var_step(O.C.byref(g), O.ptr(a), O.ptr(b), O.ptr(k_host), tau, O.BORDER_VALUE, 0.0)
    t1 = time.perf_counter() - t0
    steps = max(1, min(200, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for s in range(steps):
        L.oracle_heat7var_step(O.C.byref(g), O.ptr(a), O.ptr(b), O.ptr(k_host), tau, O.BORDER_VALUE, 0.0)
        a, b = b, a
    dt = time.perf_counter() - t0
    glups = steps * float(n) ** 3 / dt / 1e9
    return {"value": round(glups, 4), "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 variable-kappa step (restated reference CPU path, OpenMP) on the same "
                      f"{n}^3 input, {steps} steps in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--autotune", action="store_true", help="autotune tiles before timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    ensure_built()
    from masterthesis_amd import capi
    n = args.n
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / (6.0 * 1.5)
    g = capi.geom((n, n, n))
    ctx = capi.Context(0)
    cu, arch = ctx.info()
    log(f"[bench] device {arch} with {cu} CUs; workload heat7var {n}^3 x {args.steps} steps")
    ne = capi.memall_count(g, 1)
    u0 = ctx.alloc(ne, np.float64)
    k0 = ctx.alloc(ne, np.float64)
    u0.fill(0.0)
    k0.fill(0.0)
    capi.check(ctx.L.scafes_heat7var_init(ctx.h, capi.C.byref(g), 1, capi.F64, u0.ptr, k0.ptr, 0, 1e-3))
    ctx.sync()

    tune = None
    if args.autotune:
        # tune on scratch buffers, then the timed run uses the chosen tile
        ms, res, cres, bufs = run_mode(ctx, capi, g, capi.HEATVAR_F32_KAHAN, np.float32, u0, k0, tau, 2, 0)
        best, bms = capi.C.c_int(-1), capi.C.c_float(0)
        a, b, k, c1, c2 = bufs
        capi.check(ctx.L.scafes_autotune_heat7var(ctx.h, capi.C.byref(g), 1, capi.HEATVAR_F32_KAHAN, a.ptr,
                                                  b.ptr, c1.ptr, c2.ptr, k.ptr, tau, 10, capi.C.byref(best),
                                                  capi.C.byref(bms)))
        tune = {"cfg": best.value, "ms": round(bms.value, 4)}
        for d in bufs:
            d.free()
        log(f"[bench] autotune: cfg {best.value} {bms.value:.3f} ms/step")

    ms32, r32, c32, bufs32 = run_mode(ctx, capi, g, capi.HEATVAR_F32_KAHAN, np.float32, u0, k0, tau,
                                      args.steps, args.warmup, slot0=0)
    lups = float(n) ** 3 * args.steps
    glups = lups / (ms32 * 1e-3) / 1e9
    per_launch_ms = ms32 / args.steps
    achieved = BYTES_PER_LUP["f32_kahan"] * float(n) ** 3 / (per_launch_ms * 1e-3) / 1e9
    log(f"[bench] fp32+kahan: {ms32:.1f} ms, {glups:.1f} GLUPS, {achieved:.0f} GB/s")
    out = {
        "metric": METRIC, "value": round(glups, 3), "unit": "GLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per_launch_ms, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (device-seeded splitmix64 noise on the analytic sine mode)",
        "config": {"workload": f"heat7var {n}^3 nodes, variable kappa, FP32 + Kahan, {args.steps} Euler steps",
                   "nodes": n ** 3, "tau": tau, "kernel": "stream_kernel<HeatVarOp<float,2>>"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_lup": BYTES_PER_LUP["f32_kahan"]},
    }
    if tune:
        out["config"]["autotune"] = tune

    if not args.no_fp64:
        ms64, r64, _, bufs64 = run_mode(ctx, capi, g, capi.HEATVAR_F64, np.float64, u0, k0, tau,
                                        args.steps, 2, slot0=2)
        glups64 = lups / (ms64 * 1e-3) / 1e9
        ach64 = BYTES_PER_LUP["f64"] * float(n) ** 3 / (ms64 / args.steps * 1e-3) / 1e9
        ref = nodes(r64.read(), n)
        hi = nodes(r32.read(), n).astype(np.float64)
        lo = nodes(c32.read(), n).astype(np.float64)
        scale = np.abs(ref).max()
        err_k = float(np.abs(hi + lo - ref).max() / scale)
        err_hi = float(np.abs(hi - ref).max() / scale)
        del ref, hi, lo
        ms_p, r_p, _, bufs_p = run_mode(ctx, capi, g, capi.HEATVAR_F32, np.float32, u0, k0, tau,
                                        args.steps, 2, slot0=4)
        glups_p = lups / (ms_p * 1e-3) / 1e9
        ref = nodes(r64.read(), n)
        err_p = float(np.abs(nodes(r_p.read(), n).astype(np.float64) - ref).max() / scale)
        del ref
        out["fp64"] = {"value": round(glups64, 3), "unit": "GLUPS", "hbm_frac": round(ach64 / HBM_PEAK_GBS, 4)}
        out["fp32_plain"] = {"value": round(glups_p, 3), "unit": "GLUPS", "rel_err_vs_fp64": err_p}
        out["fp32_kahan_rel_err_vs_fp64"] = err_k
        out["fp32_kahan_hi_only_rel_err_vs_fp64"] = err_hi
        log(f"[bench] fp64 {glups64:.1f} GLUPS; fp32 plain {glups_p:.1f} GLUPS err {err_p:.2e}; "
            f"kahan err {err_k:.2e}")
        for d in bufs64 + bufs_p:
            if d is not None:
                d.free()

    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, OpenMP)...")
        u_host = u0.read()
        k_host = k0.read()
        out["cpu_baseline"] = cpu_baseline(u_host, k_host, tau, n)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']} GLUPS")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
