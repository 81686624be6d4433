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
CLOCK_WARMUP_MS = 150.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ensure_built():
    from masterthesis_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


SWEEP_KERNEL = {1: "stream_kernel<HeatVarOp>", 2: "heatvar_tb2_kernel", 3: "heatvar_tb3_kernel",
                4: "heatvar_tb4_kernel"}


def sweep_plan(ctx, capi, g, mode, nsteps):
    """The sweeps (steps of each, launch order) scafes_heat7var_run takes
    for nsteps with its default tile on this grid, grouped into runs of
    equal sweeps: [(steps per sweep, sweeps), ...]."""
    cnt = capi.C.c_int(0)
    capi.check(ctx.L.scafes_heat7var_plan(ctx.h, capi.C.byref(g), 1, mode, nsteps, None, 0, capi.C.byref(cnt)))
    buf = (capi.C.c_int * max(cnt.value, 1))()
    capi.check(ctx.L.scafes_heat7var_plan(ctx.h, capi.C.byref(g), 1, mode, nsteps, buf, cnt.value,
                                          capi.C.byref(cnt)))
    segs = []
    for k in buf[:cnt.value]:
        if segs and segs[-1][0] == k:
            segs[-1][1] += 1
        else:
            segs.append([k, 1])
    return [tuple(x) for x in segs]


def run_mode(ctx, capi, g, mode, dtype, u_init, k_init, tau, nsteps, warmup, sps):
    """Allocate, copy the initial state, warm up, time nsteps.

    The run is the library's default split of nsteps into sweeps
    (scafes_heat7var_plan: three-step sweeps and, where they do not divide
    nsteps, two- or four-step ones).  Warm-up (on the same buffers, then
    reset to the initial state): the requested warmup steps, then one sweep
    of every kind the timed run launches, so no kernel is launched for the
    first time inside the timed region.  The timed region is one call per
    run of equal sweeps, with an event between them; the run of sps-step
    sweeps gives the dominant kernel's live launch time.

    Returns (ms_total, ms_dominant_sweeps, dominant_sweeps, final_u, final_c,
    buffers, plan)."""
    ne = capi.memall_count(g, 1)
    code = capi.F64 if dtype == np.float64 else capi.F32
    a = ctx.alloc(ne, dtype)
    b = ctx.alloc(ne, dtype)
    k = ctx.alloc(ne, dtype)
    for d in (a, b, k):
        d.fill(0.0)
    capi.check(ctx.L.scafes_convert(ctx.h, k.ptr, code, k_init.ptr, capi.F64, ne))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.alloc(ne, np.float32), ctx.alloc(ne, np.float32)
    which = capi.C.c_int(0)
    gp = capi.C.byref(g)

    def reset():
        capi.check(ctx.L.scafes_convert(ctx.h, a.ptr, code, u_init.ptr, capi.F64, ne))
        capi.check(ctx.L.scafes_convert(ctx.h, b.ptr, code, u_init.ptr, capi.F64, ne))
        if c1 is not None:
            c1.fill(0.0)
            c2.fill(0.0)

    def run(ns, ua, ub, ca, cb):
        """ns steps from (ua, ca); returns the buffers holding the result."""
        capi.check(ctx.L.scafes_heat7var_run(ctx.h, gp, 1, mode, ua.ptr, ub.ptr, ca.ptr if ca else None,
                                             cb.ptr if cb else None, k.ptr, tau, ns, capi.BORDER_VALUE,
                                             0.0, -1, capi.C.byref(which)))
        return (ub, cb) if which.value else (ua, ca)

    segs = sweep_plan(ctx, capi, g, mode, nsteps)
    reset()
    if warmup > 0:
        run(warmup, a, b, c1, c2)
    for kk in sorted({kk for kk, _ in segs}):
        run(kk, a, b, c1, c2)  # one sweep of each kind
    # then untimed sweeps of the same kernels until the GPU has been busy for
    # CLOCK_WARMUP_MS: a short run right after an idle GPU starts below the
    # steady clocks a long run sees (measured: 1006 us per three-step sweep
    # after a 5-step warm-up, 885 after 300 steps, 869 over 1000 steps)
    busy = 0.0
    while busy < CLOCK_WARMUP_MS:
        ctx.record(3)
        run(sps * 10, a, b, c1, c2)
        ctx.record(4)
        busy += ctx.elapsed_ms(3, 4)
    reset()
    if len(segs) > 7:
        raise RuntimeError(f"sweep plan with {len(segs)} runs: more than the context's events")
    res, cres = a, c1
    ctx.sync()
    ctx.record(0)
    for q, (kk, m) in enumerate(segs):
        other = b if res is a else a
        cother = (c2 if cres is c1 else c1) if c1 is not None else None
        res, cres = run(kk * m, res, other, cres, cother)
        ctx.record(q + 1)
    ms = ctx.elapsed_ms(0, len(segs)) if segs else 0.0
    ms_dom, n_dom = 0.0, 0
    for q, (kk, m) in enumerate(segs):
        if kk == sps:
            ms_dom += ctx.elapsed_ms(q, q + 1)
            n_dom += m
    ctx.sync()
    return ms, ms_dom, n_dom, res, cres, [a, b, k, c1, c2], segs


def heatvar_kernel(ctx, capi, g, mode):
    """(configuration index, kernel name, time steps per launch) that
    scafes_heat7var_run uses by default for this grid and mode."""
    cfg = capi.C.c_int(-1)
    capi.check(ctx.L.scafes_heat7var_active_cfg(ctx.h, capi.C.byref(g), 1, mode, capi.C.byref(cfg)))
    tx, ty, zc, d, st = (capi.C.c_int() for _ in range(5))
    capi.check(ctx.L.scafes_heat7var_cfg_desc(cfg.value, capi.C.byref(tx), capi.C.byref(ty), capi.C.byref(zc),
                                              capi.C.byref(d), capi.C.byref(st)))
    t, m = {capi.HEATVAR_F64: ("double", 0), capi.HEATVAR_F32: ("float", 1),
            capi.HEATVAR_F32_KAHAN: ("float", 2)}[mode]
    if st.value >= 2:
        name = f"heatvar_tb{st.value}_kernel<{t}, {m}, {tx.value}, {ty.value}, {d.value}>"
    else:
        name = f"stream_kernel<scafes::HeatVarOp<{t}, {m}>, {t}, {tx.value}, {ty.value}, {d.value}>"
    return cfg.value, name, st.value, zc.value


def profiled_traffic(path=os.path.join(ROOT, "profiles", "heat7var_kahan_512.md")):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), and the
    average launch duration the same profile's kernel trace measured."""
    try:
        vals = {}
        for line in open(path):
            if ":" in line and line.split(":")[0] in ("traffic_bytes_per_launch", "avg_duration_ns"):
                k, v = line.split(":", 1)
                vals[k] = float(v)
        return vals.get("traffic_bytes_per_launch"), vals.get("avg_duration_ns")
    except OSError:
        return None, None


def _profile_fields(name, applies):
    """HBM bytes per launch and average launch time from a committed
    rocprofv3 summary (only when the run's shape is the profiled one)."""
    if not applies:
        return {}
    traffic, ns = profiled_traffic(os.path.join(ROOT, "profiles", name))
    if traffic is None:
        return {}
    return {"hbm_traffic_bytes_per_launch": traffic, "profile_avg_launch_us": round(ns / 1e3, 2),
            "profile": "profiles/" + name}


def nodes(arr, n):
    return arr.reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]


def host_cpu():
    """Threads the CPU baselines use and the host they run on.

    Threads: OMP_NUM_THREADS when the environment sets it (the CPU share a
    job gets beside one GPU), else every CPU this process may run on.  The
    record names the CPU model, SMT and NUMA layout from lscpu."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(int(env), avail) if env.isdigit() and int(env) > 0 else avail
    info = {"threads": threads, "cpus_available": avail,
            "thread_source": "OMP_NUM_THREADS" if env.isdigit() else "affinity mask"}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        want = {"Model name": "model", "Thread(s) per core": "smt_threads_per_core", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "NUMA node(s)": "numa_nodes", "CPU(s)": "cpus"}
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in want:
                info[want[key.strip()]] = val.strip()
    except (OSError, ValueError):
        info["lscpu"] = "unavailable"
    return info


# The reference's own OpenMP path, built from source in a container of 8
# cores, runs its 64^3 TimeLinSpaceLin heat update at 8.2 MLUP/s on one
# thread and 40 MLUP/s on eight; the oracle restates its per-node arithmetic
# in the same operation order without the per-node index and call overhead.
REFERENCE_CONTEXT = ("kind 'port': the oracle, a C restatement of the reference's FP64 CPU update in its operation "
                     "order (OpenMP); the reference's own OpenMP path measured 0.008 GLUPS on 1 thread and 0.040 "
                     "GLUPS on 8 threads (64^3 heat update, built from source on an 8-core host)")


def cpu_baseline(u_host, k_host, tau, n, budget_s=12.0):
    """The oracle's FP64 OpenMP step (restated reference CPU path) on the same
    512^3 input, a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    L = O.lib()
    host = host_cpu()
    threads = host["threads"]
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
                      f"{n}^3 input, {steps} steps in {dt:.1f} s",
            "host": host, "note": REFERENCE_CONTEXT}


def _oracle(threads=None):
    """The oracle (restated reference CPU path) for the CPU baselines."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = threads or host_cpu()["threads"]
    O.lib().oracle_set_threads(threads)
    return O, threads


def _timed(step, budget_s, cap=200):
    """Steps of `step` that fit a budget of about budget_s seconds: (count, seconds)."""
    t0 = time.perf_counter()
    step()
    t1 = time.perf_counter() - t0
    k = max(1, min(cap, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(k):
        step()
    return k, time.perf_counter() - t0


def cpu_yee(ncell=384, budget_s=5.0):
    """Oracle FP64 Yee leapfrog steps (OpenMP) on the same ncell^3 box."""
    O, threads = _oracle()
    f = O.yee_alloc(ncell)
    m = ncell // 2
    f[2, m, m, m] = 1.0
    k, dt = _timed(lambda: O.lib().oracle_yee_step(ncell, 0.5, O.ptr(f)), budget_s)
    del f
    return {"value": round(float(ncell) ** 3 * k / dt / 1e9, 4), "unit": "Gcell-steps/s", "cores": threads,
            "kind": "port", "sample": f"oracle FP64 Yee leapfrog step (OpenMP) on the same {ncell}^3 cells, "
                                      f"{k} steps in {dt:.1f} s"}


def cpu_yee_eps(ncell=128, nsteps=80):
    """Oracle FP64 permittivity gradient (forward states + reverse sweep,
    OpenMP) on a ncell^3 box of the same kind: gradient cell-steps per s."""
    O, threads = _oracle()
    m = ncell + 1
    f0 = O.yee_alloc(ncell)
    f0[2, m // 2, m // 2, m // 2] = 1.0
    ax = np.arange(m)
    r2 = (ax[:, None, None] - 0.6 * ncell) ** 2 + (ax[None, :, None] - 0.6 * ncell) ** 2 + \
        (ax[None, None, :] - 0.5 * ncell) ** 2
    eps = np.where(r2 <= (ncell / 4.0) ** 2, 4.0, 1.0)
    target = np.zeros((3, m, m, m))
    t0 = time.perf_counter()
    st = O.yee_eps_forward_all(ncell, 0.5, eps, f0, nsteps)
    O.yee_eps_grad(ncell, 0.5, eps, st, target, nsteps)
    dt = time.perf_counter() - t0
    del st
    return {"value": round(float(ncell) ** 3 * nsteps / dt / 1e9, 4), "unit": "gradient Gcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"oracle FP64 Yee eps forward (all states kept) + adjoint gradient (OpenMP), {ncell}^3 cells, "
                      f"{nsteps} steps in {dt:.1f} s"}


def cpu_pennes(n=256, budget_s=5.0):
    """Oracle FP64 Pennes forward steps (OpenMP) on the same n^3 workload."""
    from masterthesis_amd.workloads import pennes_workload
    O, threads = _oracle()
    wl = pennes_workload(n, seed=0)
    g = O.make_grid((n, n, n))
    coef = O.pennes_coef(wl["tau"], 0.5, 1.0, 1.0, 0.0)
    a = np.zeros(O.memall_shape(g))
    b = np.zeros_like(a)
    w, q = (np.ascontiguousarray(wl[k].reshape(a.shape)) for k in ("w", "q"))
    st = [a, b]

    def step():
        O.lib().oracle_pennes_step(O.C.byref(g), O.ptr(st[0]), O.ptr(st[1]), O.ptr(w), O.ptr(q), O.C.byref(coef))
        st.reverse()

    k, dt = _timed(step, budget_s)
    return {"value": round(float(n) ** 3 * k / dt / 1e9, 4), "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 Pennes forward step (OpenMP) on the same {n}^3 input, {k} steps in {dt:.1f} s"}


def cpu_batch(ens, nsteps, nsample=256):
    """Oracle FP64 2-D heat runs of the first nsample ensemble members (one
    problem per thread)."""
    from concurrent.futures import ThreadPoolExecutor
    O, threads = _oracle(1)
    L = O.lib()
    members = []
    for i in range(min(nsample, len(ens["n"]))):
        m, o = int(ens["n"][i]), int(ens["off"][i])
        u = ens["u0"][o:o + m * m].astype(np.float64)
        members.append((m, float(ens["kappa"][i]), float(ens["tau"][i]), u, np.empty_like(u)))
    pool = host_cpu()["threads"]

    def run(mem):
        m, kap, tau, u, up = mem
        L.oracle_heat2d_run(m, 1.0 / (m - 1), kap, tau, nsteps, O.ptr(u), O.ptr(up))

    t0 = time.perf_counter()
    with ThreadPoolExecutor(pool) as ex:
        list(ex.map(run, members))
    dt = time.perf_counter() - t0
    lups = sum(m * m for m, *_ in members) * float(nsteps)
    return {"value": round(lups / dt / 1e9, 4), "unit": "GLUPS", "cores": pool, "kind": "port",
            "sample": f"oracle FP64 2-D heat runs ({nsteps} steps) of the first {len(members)} of the "
                      f"{len(ens['n'])} problems, one problem per thread, {dt:.1f} s"}


def cpu_emfdtd(n=256, budget_s=6.0):
    """Oracle FP64 EMFDTD steps (serial) on the same n^3 case."""
    from masterthesis_amd.workloads import sphere_sigma
    O, _ = _oracle()
    ref = O.Emfdtd(n, d=10, src=(25, 25, 25), sink=(30, 30, 30), sigma=sphere_sigma(n, (n / 2, n / 2, n / 2), 50.0))
    k, dt = _timed(lambda: ref.step(0.0), budget_s, cap=20)
    del ref
    return {"value": round(float(n) ** 3 * k / dt / 1e9, 4), "unit": "Gnode-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 EMFDTD leapfrog step (serial, as the reference's hooks run per rank) on the "
                      f"same {n}^3 case, {k} steps in {dt:.1f} s"}


def bench_heat64(ctx, capi, n=64, nsteps=100):
    """HeatEqnFDM plain method: u_t = Laplace(u) on the unit cube, n^3 nodes,
    FP64, tau = 0.9 h^2/6, zero Dirichlet border, u0 = prod sin(pi x_i):
    GLUPS, the error against the analytic solution exp(-3 pi^2 t) prod
    sin(pi x_i), and against the discrete scheme's closed form
    (1 - tau lambda_h)^n prod sin(pi x_i)."""
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 6.0
    x = np.linspace(0.0, 1.0, n)
    s = np.sin(np.pi * x)
    mode = s[None, None, :] * s[None, :, None] * s[:, None, None]
    u0 = np.zeros((n + 2,) * 3)
    u0[1:-1, 1:-1, 1:-1] = mode
    g = capi.geom((n, n, n))
    p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, border_mode=capi.BORDER_VALUE)
    a, b = ctx.upload(u0), ctx.upload(u0)
    gp = capi.C.byref(g)

    which = capi.C.c_int(0)

    def run(k, x0, x1):
        capi.check(ctx.L.scafes_heat7_run_f64(ctx.h, gp, 1, x0.ptr, x1.ptr, k, capi.C.byref(p),
                                              capi.C.byref(which)))
        return x1 if which.value else x0

    run(10, a, b)
    a.write(u0)
    b.write(u0)
    ctx.sync()
    ctx.record(0)
    res = run(nsteps, a, b)
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1)
    u = res.read().reshape(u0.shape)[1:-1, 1:-1, 1:-1]
    a.free()
    b.free()
    t = nsteps * tau
    lam_h = 3 * (4.0 / (h * h)) * np.sin(np.pi * h / 2) ** 2
    exact_discrete = (1.0 - tau * lam_h) ** nsteps * mode
    analytic = np.exp(-3 * np.pi ** 2 * t) * mode
    return {"workload": f"heat {n}^3 FP64, sine mode, zero Dirichlet, tau = 0.9 h^2/6, {nsteps} steps",
            "kernel": "stream_kernel<HeatConstOp> (one step per launch, scafes_heat7_run_f64)", "ms": round(ms, 4),
            "GLUPS": round(float(n) ** 3 * nsteps / (ms * 1e-3) / 1e9, 3),
            "rel_err_vs_discrete_closed_form": float(np.abs(u - exact_discrete).max() / np.abs(exact_discrete).max()),
            "rel_err_vs_analytic": float(np.abs(u - analytic).max() / np.abs(analytic).max())}


def cpu_heat64(n=64, budget_s=3.0):
    """Oracle FP64 constant-coefficient heat steps (OpenMP) on the same n^3 case."""
    O, threads = _oracle()
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    coef = O.heat7_coef(0.9 * h * h / 6.0)
    a = np.zeros(O.memall_shape(g))
    x = np.linspace(0.0, 1.0, n)
    s = np.sin(np.pi * x)
    O.nodes_view(g, a)[...] = s[None, None, :] * s[None, :, None] * s[:, None, None]
    b = a.copy()
    st = [a, b]

    def step():
        O.lib().oracle_heat7_step(O.C.byref(g), O.ptr(st[0]), O.ptr(st[1]), None, None, None, O.BORDER_VALUE, 0.0,
                                  O.C.byref(coef))
        st.reverse()

    k, dt = _timed(step, budget_s, cap=2000)
    return {"value": round(float(n) ** 3 * k / dt / 1e9, 4), "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 heat step (OpenMP) on the same {n}^3 case, {k} steps in {dt:.1f} s"}


def bench_yee(ctx, capi, ncell=384, nsteps=2000, energy_every=100):
    """Yee FDTD, PEC box of ncell^3 cells, vacuum, Courant 0.5, Gaussian Ez
    pulse (sigma = 8 cells) at the centre, FP32 fields, FP64 energy."""
    nf = (ncell + 1) ** 3
    f = ctx.alloc(6 * nf, np.float32)
    f2 = ctx.alloc(6 * nf, np.float32)  # ping-pong state: one fused sweep per step
    second = capi.C.c_int(0)
    capi.check(ctx.L.scafes_yee_init_pulse(ctx.h, ncell, capi.F32, f.ptr, 8.0))
    capi.check(ctx.L.scafes_yee_run(ctx.h, ncell, capi.F32, 0.5, f.ptr, f2.ptr, 2, 0, None, None))
    capi.check(ctx.L.scafes_yee_init_pulse(ctx.h, ncell, capi.F32, f.ptr, 8.0))
    nrec = (nsteps + energy_every - 1) // energy_every
    e = np.zeros(nrec)
    ctx.sync()
    ctx.record(0)
    capi.check(ctx.L.scafes_yee_run(ctx.h, ncell, capi.F32, 0.5, f.ptr, f2.ptr, nsteps, energy_every,
                                    e.ctypes.data_as(capi.C.POINTER(capi.C.c_double)), capi.C.byref(second)))
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1)
    f.free()
    f2.free()
    cells = float(ncell) ** 3
    per = ms / nsteps
    return {"workload": f"yee {ncell}^3 cells PEC vacuum, Courant 0.5, {nsteps} leapfrog steps, FP32",
            "kernel": "yee_tb2_kernel<float, 64, 16, 4> (two leapfrog steps per sweep)",
            "ms_per_step": round(per, 5), "gcell_steps_per_s": round(cells / (per * 1e-3) / 1e9, 3),
            "algorithmic_bytes_per_cell_step": 48, "moved_bytes_per_cell_step": 24,
            "hbm_frac_algorithmic": round(48 * cells / (per * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "energy_rel_drift_fp64": float(np.abs(e - e[0]).max() / e[0]), "energy_samples": int(nrec),
            **_profile_fields("yee_tb2_384.md", ncell == 384)}


def bench_kappa_grad(ctx, capi, g, n, u0, k0, tau, nsteps, t_fwd_ms, snapshots=(30, 100)):
    """dJ/dkappa of the headline problem (the same n^3 input, kappa field,
    tau and zero Dirichlet border; FP32+Kahan forward, FP32 adjoint field,
    FP64 gradient), J = 0.5 sum_inner (u^N - u^0)^2, for two snapshot
    budgets, against the headline forward time of the same nsteps."""
    ne = capi.memall_count(g, 1)
    uf, kf = ctx.alloc(ne, np.float32), ctx.alloc(ne, np.float32)
    capi.check(ctx.L.scafes_convert(ctx.h, uf.ptr, capi.F32, u0.ptr, capi.F64, ne))
    capi.check(ctx.L.scafes_convert(ctx.h, kf.ptr, capi.F32, k0.ptr, capi.F64, ne))
    grad = ctx.alloc(ne, np.float64)
    out = {"workload": f"heat7var {n}^3 dJ/dkappa, {nsteps} steps, FP32+Kahan forward, FP32 adjoint, FP64 gradient",
           "fwd_ms": round(t_fwd_ms, 2),
           "recompute": "the headline run's sweeps where the schedule advances several steps (three-step "
                        "sweeps, four-step ones where a count is not a multiple of three); where it stores "
                        "every state, the same sweeps writing the states in between "
                        "(heatvar_tb3_kernel<float, 2, 64, 12, 2, true>, 28 B/node per sweep; "
                        "heatvar_tb4_kernel<float, 2, 64, 16, 2, true>, 32 B/node per sweep); else one step",
           "adjoint_kernels": "heatvar_adjk_kernel<float, 64, 16, K, 2> (K <= 4 reverse steps per sweep where the "
                              "schedule has the states at hand, (28 + 4K) B/node per sweep: lam r+w, kappa, "
                              "u^{n-1..n-K}, FP64 gradient r+w); stream_kernel<HeatVarAdjOp<float>> (one step, "
                              "32 B) elsewhere",
           "adjoint_profile": _profile_fields("heat7var_adjoint_512.md", n == 512),
           "gradient": []}
    for s in snapshots:
        J = capi.C.c_double(0)
        stats = np.zeros(4)

        def call():
            return ctx.L.scafes_heat7var_gradient(ctx.h, capi.C.byref(g), 1, capi.HEATVAR_F32_KAHAN, uf.ptr, kf.ptr,
                                                  uf.ptr, tau, capi.BORDER_VALUE, 0.0, nsteps, s, grad.ptr,
                                                  capi.C.byref(J),
                                                  stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double)))
        rc, t_first, t_grad = _grad_calls(ctx, call)
        if rc != 0:
            out["gradient"].append({"snapshots": s, "error": ctx.L.scafes_last_error().decode()})
            continue
        out["gradient"].append({"snapshots": int(stats[2]), "checkpoint_GB": round(float(stats[3]) / 1e9, 2),
                                "forward_steps": int(stats[0]), "adjoint_steps": int(stats[1]),
                                "grad_ms": round(t_grad, 2), "grad_first_call_ms": round(t_first, 2),
                                "grad_over_fwd": round(t_grad / t_fwd_ms, 3), "J": J.value})
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))
    for d in (uf, kf, grad):
        d.free()
    return out


def _grad_calls(ctx, call):
    """A gradient entry point called twice: (status, first call ms, second
    call ms).  The first call sizes the workspace the context keeps (fresh
    allocations, or memory the driver must clear after an earlier free); the
    second is the cost of each further gradient, e.g. in an optimisation loop."""
    times = []
    for _ in range(2):
        ctx.sync()
        t0 = time.perf_counter()
        rc = call()
        times.append((time.perf_counter() - t0) * 1e3)
        if rc != 0:
            return rc, 0.0, 0.0
    return 0, times[0], times[1]


def bench_yee_eps(ctx, capi, ncell=384, nsteps=200, snapshots=(20, 100)):
    """Yee FDTD with a permittivity field: PEC box of ncell^3 cells, Courant
    0.5, Gaussian Ez pulse (sigma = 8 cells) at the centre, a dielectric
    sphere (radius ncell/4, eps = 4) off centre; forward time (eps sweeps and
    the vacuum two-steps-per-sweep kernel) vs the adjoint gradient of the
    final-E misfit against the E field of a sphere with eps = 4.4 (FP32
    fields, FP64 gradient) for two snapshot budgets."""
    m = ncell + 1
    nf = m ** 3
    dt = np.float32
    ax = np.arange(m, dtype=np.float32)
    c0 = 0.6 * ncell
    r2 = (ax[:, None, None] - c0) ** 2 + (ax[None, :, None] - c0) ** 2 + (ax[None, None, :] - 0.5 * ncell) ** 2
    inside = r2 <= (ncell / 4.0) ** 2
    eps = ctx.upload(np.where(inside, 4.0, 1.0).astype(dt))
    eps_t = ctx.upload(np.where(inside, 4.4, 1.0).astype(dt))
    del r2, inside
    f0 = ctx.alloc(6 * nf, dt)
    capi.check(ctx.L.scafes_yee_init_pulse(ctx.h, ncell, capi.F32, f0.ptr, 8.0))
    f, w = ctx.alloc(6 * nf, dt), ctx.alloc(6 * nf, dt)

    def eps_forward(ebuf, steps):
        capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, f.ptr, f0.ptr, 6 * nf * 4))
        capi.check(ctx.L.scafes_yee_eps_run(ctx.h, ncell, capi.F32, 0.5, ebuf.ptr, f.ptr, w.ptr, steps))

    eps_forward(eps_t, nsteps)
    target = ctx.alloc(3 * nf, dt)
    capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, target.ptr, f.ptr, 3 * nf * 4))
    eps_t.free()
    eps_forward(eps, 4)
    ctx.sync()
    ctx.record(0)
    eps_forward(eps, nsteps)
    ctx.record(1)
    t_eps = ctx.elapsed_ms(0, 1)
    capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, f.ptr, f0.ptr, 6 * nf * 4))
    ctx.sync()
    ctx.record(0)
    capi.check(ctx.L.scafes_yee_run(ctx.h, ncell, capi.F32, 0.5, f.ptr, w.ptr, nsteps, 0, None, None))
    ctx.record(1)
    t_vac = ctx.elapsed_ms(0, 1)
    cells = float(ncell) ** 3
    out = {"workload": f"yee {ncell}^3 cells PEC, dielectric sphere eps=4 (r={ncell // 4}), Courant 0.5, "
                       f"{nsteps} steps, FP32 fields, FP64 gradient",
           "fwd_eps_ms": round(t_eps, 2), "fwd_eps_gcell_steps_per_s": round(cells * nsteps / (t_eps * 1e-3) / 1e9, 2),
           "fwd_vacuum_ms": round(t_vac, 2),
           "fwd_kernels": "yee_tb2_kernel<float, 64, 16, 4, true> (two steps per sweep, 26 B/cell-step algorithmic); "
                          "yee_fused_kernel<float, 64, 8, 2, true> for a single step (52 B); in runs of one-step "
                          "binomial splits the two-step sweep that also writes the state after one step "
                          "(yee_tb2_kernel<float, 64, 16, 4, true, true>)",
           "adjoint_kernel": "adj_fused<float, 64, 16, 2> (one reverse step per sweep, 96 B/cell-step algorithmic)",
           **({"adjoint_profile": _profile_fields("yee_eps_adjoint_384.md", True)} if ncell == 384 else {}),
           "gradient": []}
    grad = ctx.alloc(nf, np.float64)
    for s in snapshots:
        J = capi.C.c_double(0)
        stats = np.zeros(4)

        def call():
            return ctx.L.scafes_yee_eps_gradient(ctx.h, ncell, capi.F32, 0.5, f0.ptr, eps.ptr, target.ptr, nsteps,
                                                 s, grad.ptr, capi.C.byref(J),
                                                 stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double)))
        rc, t_first, t_grad = _grad_calls(ctx, call)
        if rc != 0:
            out["gradient"].append({"snapshots": s, "error": ctx.L.scafes_last_error().decode()})
            continue
        out["gradient"].append({"snapshots": int(stats[2]), "checkpoint_GB": round(float(stats[3]) / 1e9, 2),
                                "forward_steps": int(stats[0]), "adjoint_steps": int(stats[1]),
                                "grad_ms": round(t_grad, 2), "grad_first_call_ms": round(t_first, 2),
                                "grad_gcell_steps_per_s": round(cells * nsteps / (t_grad * 1e-3) / 1e9, 3),
                                "grad_over_fwd_eps": round(t_grad / t_eps, 3),
                                "grad_over_fwd_vacuum": round(t_grad / t_vac, 3), "J": J.value})
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))
    for d in (eps, f0, f, w, target, grad):
        d.free()
    return out


# (first index, last index - (n-1), axis of the l term, axis of the r term)
# of the six curl stencils (EMFDTD.hpp:167-172, 208-213)
_EM_STENCILS = [((1, 0, 0), (0, -1, -1), 2, 1), ((0, 1, 0), (-1, 0, -1), 0, 2), ((0, 0, 1), (-1, -1, 0), 1, 0),
                ((0, 0, 1), (-1, 0, 0), 1, 2), ((1, 0, 1), (0, -1, 0), 2, 0), ((1, 1, 0), (0, 0, -1), 0, 1)]


def _emfdtd_psi_terms(n, d):
    """Number of (node, psi term) pairs inside the PML per step: a term is
    in the PML where the node is in its stencil's range and its distance to
    the range's nearer end along the term's axis is below d."""
    total = 0
    for lo, end, al, ar in _EM_STENCILS:
        hi = [n - 1 + e for e in end]
        cnt = [hi[q] - lo[q] + 1 for q in range(3)]
        for ax in (al, ar):
            idx = np.arange(lo[ax], hi[ax] + 1)
            pml = int(np.count_nonzero(np.minimum(idx, hi[ax] - idx) < d))
            total += pml * cnt[(ax + 1) % 3] * cnt[(ax + 2) % 3]
    return float(total)


def bench_emfdtd(ctx, capi, n=256, nsteps=500):
    """EMFDTD example (Yee leapfrog + complete CPML + conductive sphere) on
    n^3 nodes, FP64: the [PML] / [source] / [sink] parameters of
    examples/EMFDTD/testcase_normal.ini (d = 10, source 25x25x25, sink
    30x30x30, unmodulated pulse t0 = 40, spread = 8) with a sphere of radius
    50 at the centre."""
    from masterthesis_amd.workloads import sphere_sigma
    d = 10
    p = capi.emfdtd_params(n, d=d, src=(25, 25, 25), sink=(30, 30, 30))
    nn = n ** 3
    sigma = sphere_sigma(n, (n / 2, n / 2, n / 2), 50.0).ravel()
    ga, gb = capi.emfdtd_ga_gb(sigma, p.dt)
    state = ctx.alloc(capi.EM_NFIELDS * nn, np.float64)
    state.fill(0.0)
    vp = capi.C.c_void_p
    for q, arr in ((24, ga), (25, gb)):
        capi.check(ctx.L.scafes_memcpy_h2d(ctx.h, vp(state.ptr.value + q * nn * 8), arr.ctypes.data_as(vp), nn * 8))
    sink = ctx.alloc(nsteps, np.float64)
    capi.check(ctx.L.scafes_emfdtd_run(ctx.h, capi.C.byref(p), state.ptr, 2, 0.0, sink.ptr))  # warm-up
    state.fill(0.0)
    for q, arr in ((24, ga), (25, gb)):
        capi.check(ctx.L.scafes_memcpy_h2d(ctx.h, vp(state.ptr.value + q * nn * 8), arr.ctypes.data_as(vp), nn * 8))
    ctx.sync()
    ctx.record(0)
    capi.check(ctx.L.scafes_emfdtd_run(ctx.h, capi.C.byref(p), state.ptr, nsteps, 0.0, sink.ptr))
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1)
    sk = sink.read()
    state.free()
    sink.free()
    # algorithmic bytes per node-step: H phase reads H, E and writes H (72 B);
    # E phase reads D, H, gb and writes D, E (104 B) in vacuum, plus ga and I
    # read / I written (56 B) in the conductor; each PML psi term adds 16 B
    cond = float(np.count_nonzero(gb != 0.0))
    alg = 176.0 + 56.0 * cond / nn + 16.0 * _emfdtd_psi_terms(n, d) / nn
    per = ms / nsteps
    rate = nn / (per * 1e-3)
    return {"workload": f"EMFDTD {n}^3 nodes, CPML d={d}, conductive sphere r=50, {nsteps} leapfrog steps, FP64",
            "kernels": "em_h_kernel + em_e_kernel (one pass each per step)",
            "ms_per_step": round(per, 4), "gnode_steps_per_s": round(rate / 1e9, 3),
            "algorithmic_bytes_per_node_step": round(alg, 1),
            "hbm_frac_algorithmic": round(alg * rate / 1e9 / HBM_PEAK_GBS, 4),
            "sink_peak": float(np.abs(sk).max())}


def bench_pennes(ctx, capi, n=256, nsteps=2000, snapshots=(62, 500, 1999)):
    """Pennes bioheat on n^3: k = 0.5, rho c = 1, T_a = 37, log-normal
    perfusion w (median 0.5, sigma 0.3, i.i.d. seeded), Gaussian source;
    forward time vs adjoint-gradient time (FP32 fields, FP64 gradient) for
    binomial snapshot budgets of 62 states (4.3 GB) and 500 states (34 GB,
    an eighth of one MI355X's 288 GB) and for store-all (1999 states,
    137 GB)."""
    from masterthesis_amd.workloads import pennes_workload
    wl = pennes_workload(n, seed=0)
    g = capi.geom((n, n, n))
    ne = capi.memall_count(g, 1)
    dt = np.float32
    th0 = ctx.upload(np.zeros(ne, dtype=dt))  # theta = T - T_a, T0 = T_a
    w = ctx.upload(wl["w"].astype(dt))
    q = ctx.upload(wl["q"].astype(dt))
    pp = capi.PennesParams(tau=wl["tau"], cdiff=0.5, cperf=1.0, csrc=1.0, t_a=0.0)
    a, b = ctx.alloc(ne, dt), ctx.alloc(ne, dt)
    gp = capi.C.byref(g)

    def forward(wbuf, steps, prm):
        capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, a.ptr, th0.ptr, ne * 4))
        capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, b.ptr, th0.ptr, ne * 4))
        x, y = a, b
        for _ in range(steps):
            capi.check(ctx.L.scafes_pennes_step(ctx.h, gp, 1, capi.F32, x.ptr, y.ptr, wbuf.ptr, q.ptr,
                                                capi.C.byref(prm)))
            x, y = y, x
        return x

    # target from the perturbed perfusion
    wp = ctx.upload((wl["w"] * wl["pert"]).astype(dt))
    tgt_src = forward(wp, nsteps, pp)
    target = ctx.alloc(ne, dt)
    capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, target.ptr, tgt_src.ptr, ne * 4))
    wp.free()
    forward(w, 5, pp)
    ctx.sync()
    ctx.record(0)
    forward(w, nsteps, pp)
    ctx.record(1)
    t_fwd = ctx.elapsed_ms(0, 1)
    out = {"workload": f"pennes {n}^3, {nsteps} steps, FP32 fields (deviation from T_a), FP64 gradient",
           "fwd_ms": round(t_fwd, 2), "fwd_GLUPS": round(float(n) ** 3 * nsteps / (t_fwd * 1e-3) / 1e9, 2),
           "fwd_kernel": "stream_kernel<PennesOp<float>> (16 B/LUP)",
           "adjoint_kernel": "pennes_adjk_kernel<float> (up to four reverse steps per sweep, 11 B/step at four)",
           "fwd_profile": _profile_fields("pennes_fwd_256.md", n == 256),
           "adjoint_profile": _profile_fields("pennes_adjk_256.md", n == 256), "gradient": []}
    grad = ctx.alloc(ne, np.float64)
    for s in snapshots:
        J = capi.C.c_double(0)
        stats = np.zeros(4)

        def call():
            return ctx.L.scafes_pennes_gradient(ctx.h, gp, 1, capi.F32, th0.ptr, w.ptr, q.ptr, target.ptr,
                                                capi.C.byref(pp), nsteps, s, grad.ptr, capi.C.byref(J),
                                                stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double)))
        rc, t_first, t_grad = _grad_calls(ctx, call)
        if rc != 0:
            out["gradient"].append({"snapshots": s, "error": ctx.L.scafes_last_error().decode()})
            continue
        out["gradient"].append({"snapshots": int(stats[2]), "checkpoint_GB": round(float(stats[3]) / 1e9, 2),
                                "forward_steps": int(stats[0]), "adjoint_steps": int(stats[1]),
                                "grad_ms": round(t_grad, 2), "grad_over_fwd": round(t_grad / t_fwd, 3),
                                "grad_first_call_ms": round(t_first, 2), "J": J.value})
    # The Tumor example's borders (Neumann flux on every face, Cauchy
    # convection on the far z face; TumorHeatEqnFDM.hpp:272-349) on the same
    # workload: k_cond = cdiff (rho c = 1), H = 10, ambient 15 K below T_a,
    # no flux.  H adds 2 H / h_z = 2.6 % to the explicit-step rate, inside the
    # workload's 10 % margin.  J and the gradient cover every node.
    pt = capi.PennesParams(tau=wl["tau"], cdiff=0.5, cperf=1.0, csrc=1.0, t_a=0.0)
    pt.border_mode, pt.k_cond, pt.h_conv, pt.t_inf, pt.q_bc = capi.BORDER_TUMOR, 0.5, 10.0, -15.0, 0.0
    wp = ctx.upload((wl["w"] * wl["pert"]).astype(dt))
    tgt_src = forward(wp, nsteps, pt)
    capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, target.ptr, tgt_src.ptr, ne * 4))
    wp.free()
    forward(w, 5, pt)
    ctx.sync()
    ctx.record(0)
    forward(w, nsteps, pt)
    ctx.record(1)
    t_fwd_t = ctx.elapsed_ms(0, 1)
    J = capi.C.c_double(0)
    stats = np.zeros(4)

    def call_t():
        return ctx.L.scafes_pennes_gradient(ctx.h, gp, 1, capi.F32, th0.ptr, w.ptr, q.ptr, target.ptr,
                                            capi.C.byref(pt), nsteps, 500, grad.ptr, capi.C.byref(J),
                                            stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double)))
    rc, t_first, t_grad = _grad_calls(ctx, call_t)
    tum = {"border": "Tumor example: Neumann (q_bc = 0) on every face, Cauchy (H = 10, T_inf = T_a - 15) on z = n-1",
           "fwd_ms": round(t_fwd_t, 2)}
    if rc != 0:
        tum["error"] = ctx.L.scafes_last_error().decode()
    else:
        tum.update({"snapshots": int(stats[2]), "checkpoint_GB": round(float(stats[3]) / 1e9, 2),
                    "forward_steps": int(stats[0]), "adjoint_steps": int(stats[1]), "grad_ms": round(t_grad, 2),
                    "grad_over_fwd": round(t_grad / t_fwd_t, 3), "grad_first_call_ms": round(t_first, 2),
                    "J": J.value})
    out["tumor_border"] = tum
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))
    for d in (th0, w, q, a, b, target, grad):
        d.free()
    return out


def bench_batch(ctx, capi, nprob=4096, nsteps=500):
    """4096 2-D heat problems, n ~ U{32..256}, kappa ~ U[0.1,1], 500 FP32
    steps each plus dJ/dkappa, one persistent hipGraph-captured launch."""
    from masterthesis_amd.workloads import batch2d_ensemble
    ens = batch2d_ensemble(nprob, nsteps, seed=0)
    d = {k: ctx.upload(ens[k]) for k in ("n", "off", "kappa", "tau", "order", "target")}
    u = ctx.upload(ens["u0"])
    J = ctx.alloc(nprob, np.float64)
    gr = ctx.alloc(nprob, np.float64)

    def run(graph):
        capi.check(ctx.L.scafes_batch2d_run(ctx.h, nprob, d["n"].ptr, d["off"].ptr, d["kappa"].ptr, d["tau"].ptr,
                                            d["order"].ptr, u.ptr, d["target"].ptr, J.ptr, gr.ptr, nsteps, graph))

    run(0)
    # the first graph call checks the sizes, captures and instantiates the
    # launch and runs it; later calls with the same arguments replay the
    # kept graph.  Host wall time around each call and its completion; the
    # median of three replays is the reported time.
    ts = []
    for _ in range(4):
        u.write(ens["u0"])
        ctx.sync()
        t0 = time.perf_counter()
        run(1)
        ctx.sync()
        ts.append((time.perf_counter() - t0) * 1e3)
    first, ms = ts[0], sorted(ts[1:])[1]
    nodes = float(np.sum(ens["n"].astype(np.int64) ** 2))
    gsum = float(np.sum(np.abs(gr.read())))
    for x in list(d.values()) + [u, J, gr]:
        x.free()
    return {"workload": f"{nprob} 2-D heat problems n~U{{32..256}}, {nsteps} steps + dJ/dkappa, FP32, one launch",
            "ms": round(ms, 2), "first_call_ms": round(first, 2),
            "GLUPS": round(nodes * nsteps / (ms * 1e-3) / 1e9, 2), "nodes": int(nodes),
            "sum_abs_grad": gsum}


def calibrate_copy(ctx, capi, nbytes=1 << 30, reps=3):
    """Plain copy of known size (reads nbytes, writes nbytes) for scaling the
    HBM counters of a profile."""
    src, dst = ctx.alloc(nbytes // 4, np.float32), ctx.alloc(nbytes // 4, np.float32)
    src.fill(1.0)
    ctx.sync()
    ctx.record(0)
    for _ in range(reps):
        capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, dst.ptr, src.ptr, nbytes))
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1) / reps
    src.free()
    dst.free()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(ms, 4),
            "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--autotune", action="store_true", help="autotune tiles before timing")
    ap.add_argument("--no-yee", action="store_true", help="skip the Yee FDTD leg")
    ap.add_argument("--no-pennes", action="store_true", help="skip the Pennes adjoint leg")
    ap.add_argument("--no-yee-eps", action="store_true", help="skip the Yee permittivity-gradient leg")
    ap.add_argument("--no-kappa-grad", action="store_true", help="skip the headline kappa-gradient leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-emfdtd", action="store_true", help="skip the EMFDTD (CPML) leg")
    ap.add_argument("--calibrate-copy", action="store_true", help="also run a 1 GiB copy (counter calibration)")
    ap.add_argument("--leg-steps", type=int, default=0,
                    help="step count of the Yee / Pennes / batch legs (0 = their workloads' counts; "
                         "short runs are for profiling)")
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
        # tune on scratch buffers; the timed run then uses the choice kept for this grid
        ne = capi.memall_count(g, 1)
        bufs = [ctx.alloc(ne, np.float32) for _ in range(5)]
        a, b, k, c1, c2 = bufs
        for d in (a, b, c1, c2):
            d.fill(0.0)
        capi.check(ctx.L.scafes_convert(ctx.h, a.ptr, capi.F32, u0.ptr, capi.F64, ne))
        capi.check(ctx.L.scafes_convert(ctx.h, k.ptr, capi.F32, k0.ptr, capi.F64, ne))
        best, bms = capi.C.c_int(-1), capi.C.c_float(0)
        capi.check(ctx.L.scafes_autotune_heat7var(ctx.h, capi.C.byref(g), 1, capi.HEATVAR_F32_KAHAN, a.ptr,
                                                  b.ptr, c1.ptr, c2.ptr, k.ptr, tau, 12, capi.C.byref(best),
                                                  capi.C.byref(bms)))
        tune = {"cfg": best.value, "ms_per_step": round(bms.value, 4)}
        for d in bufs:
            d.free()
        log(f"[bench] autotune: cfg {best.value} {bms.value:.3f} ms/step")

    cfg, kname, sps, zc = heatvar_kernel(ctx, capi, g, capi.HEATVAR_F32_KAHAN)
    ms32, ms32_full, launches, r32, c32, bufs32, plan32 = run_mode(ctx, capi, g, capi.HEATVAR_F32_KAHAN,
                                                                   np.float32, u0, k0, tau, args.steps,
                                                                   args.warmup, sps)
    lups = float(n) ** 3 * args.steps
    glups = lups / (ms32 * 1e-3) / 1e9
    B = BYTES_PER_LUP["f32_kahan"]
    # algorithmic-equivalent rate: every LUP of the timed region charged its
    # 20 B (u, c, kappa read; u, c written), as the metric defines HBM %
    achieved = B * lups / (ms32 * 1e-3) / 1e9
    # the dominant kernel's live launch time (its sweeps in the timed region)
    # and the bytes one sweep must move (each field read once, written once)
    launch_us = ms32_full * 1e3 / launches if launches else None
    sweep_bytes = B * float(n) ** 3
    copy = calibrate_copy(ctx, capi)
    # the box's best plain copy through the library: 4 GiB each way (larger
    # copies amortise the launch's ramp-up and tail)
    copy4 = calibrate_copy(ctx, capi, nbytes=4 << 30)
    log(f"[bench] fp32+kahan ({kname}): {ms32:.2f} ms, {glups:.1f} GLUPS, {achieved:.0f} GB/s algorithmic-equivalent; "
        f"{launches} sweeps of {sps} steps at {launch_us} us; sweeps {plan32}")
    out = {
        "metric": METRIC, "value": round(glups, 3), "unit": "GLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms32 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (device-seeded splitmix64 noise on the analytic sine mode)",
        "config": {"workload": f"heat7var {n}^3 nodes, variable kappa, FP32 + Kahan, {args.steps} Euler steps",
                   "nodes": n ** 3, "tau": tau, "kernel": kname, "cfg": cfg, "steps_per_launch": sps,
                   "z_planes_per_workgroup": zc,
                   "sweeps": ", ".join(f"{m} x {kk} steps ({SWEEP_KERNEL[kk]})" for kk, m in plan32),
                   "warmup_note": f"{args.warmup} warm-up steps, one sweep of each kind the timed run launches, "
                                  f"then untimed sweeps of the same kernels until the GPU has been busy "
                                  f"{CLOCK_WARMUP_MS:.0f} ms (steady clocks); the state is reset before timing"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "accounting": "algorithmic-equivalent: 20 B per LUP times the LUPs of the timed region, "
                                   "divided by its duration (HIP events on the library stream); a sweep of "
                                   f"{sps} steps moves one sweep's bytes for {sps} LUPs per node, so this rate "
                                   "is not HBM traffic: frac_hbm_measured is",
                     "algorithmic_bytes_per_lup": B, "timed_lups": lups,
                     "dominant_kernel": kname, "steps_per_launch": sps,
                     "avg_launch_us": round(launch_us, 2) if launch_us else None,
                     "one_sweep_bytes": sweep_bytes,
                     "frac_one_sweep": round(sweep_bytes / (launch_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                     if launch_us else None,
                     "copy_GBps": copy["GBps"], "copy_4GiB_GBps": copy4["GBps"]},
    }
    traffic, prof_ns = profiled_traffic()
    if traffic is not None and n == N:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["profile_avg_launch_us"] = round(prof_ns / 1e3, 2)
        out["roofline"]["frac_hbm_measured"] = round(traffic / (prof_ns * 1e-9) / 1e9 / HBM_PEAK_GBS, 4)
        # the same traffic rate against the best copy this run measured
        out["roofline"]["frac_of_copy_4GiB"] = round(traffic / (prof_ns * 1e-9) / 1e9 / copy4["GBps"], 4)
        out["roofline"]["profile"] = "profiles/heat7var_kahan_512.md"
    if tune:
        out["config"]["autotune"] = tune

    if not args.no_fp64:
        _, kname64, sps64, _ = heatvar_kernel(ctx, capi, g, capi.HEATVAR_F64)
        ms64, _, _, r64, _, bufs64, _ = run_mode(ctx, capi, g, capi.HEATVAR_F64, np.float64, u0, k0, tau,
                                              args.steps, args.warmup, sps64)
        glups64 = lups / (ms64 * 1e-3) / 1e9
        ref = nodes(r64.read(), n)
        hi = nodes(r32.read(), n).astype(np.float64)
        lo = nodes(c32.read(), n).astype(np.float64)
        scale = np.abs(ref).max()
        err_k = float(np.abs(hi + lo - ref).max() / scale)
        err_hi = float(np.abs(hi - ref).max() / scale)
        del ref, hi, lo
        _, kname_p, sps_p, _ = heatvar_kernel(ctx, capi, g, capi.HEATVAR_F32)
        ms_p, _, _, r_p, _, bufs_p, _ = run_mode(ctx, capi, g, capi.HEATVAR_F32, np.float32, u0, k0, tau,
                                              args.steps, args.warmup, sps_p)
        glups_p = lups / (ms_p * 1e-3) / 1e9
        ref = nodes(r64.read(), n)
        err_p = float(np.abs(nodes(r_p.read(), n).astype(np.float64) - ref).max() / scale)
        del ref
        out["fp64"] = {"value": round(glups64, 3), "unit": "GLUPS", "kernel": kname64,
                       "hbm_frac_algorithmic_equivalent": round(BYTES_PER_LUP["f64"] * glups64 / HBM_PEAK_GBS, 4)}
        out["fp32_plain"] = {"value": round(glups_p, 3), "unit": "GLUPS", "kernel": kname_p,
                             "rel_err_vs_fp64": err_p,
                             "hbm_frac_algorithmic_equivalent": round(BYTES_PER_LUP["f32"] * glups_p / HBM_PEAK_GBS,
                                                                      4)}
        out["fp32_kahan_rel_err_vs_fp64"] = err_k
        out["fp32_kahan_hi_only_rel_err_vs_fp64"] = err_hi
        log(f"[bench] fp64 {glups64:.1f} GLUPS; fp32 plain {glups_p:.1f} GLUPS err {err_p:.2e}; "
            f"kahan err {err_k:.2e}")
        for d in bufs64 + bufs_p:
            if d is not None:
                d.free()

    for d in bufs32:
        if d is not None:
            d.free()
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, OpenMP)...")
        u_host = u0.read()
        k_host = k0.read()
        out["cpu_baseline"] = cpu_baseline(u_host, k_host, tau, n)
        del u_host, k_host
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']} GLUPS")
    if not args.no_kappa_grad:
        out["kappa_gradient"] = bench_kappa_grad(ctx, capi, g, n, u0, k0, tau, args.steps, ms32)
        log(f"[bench] kappa gradient {out['kappa_gradient']}")
    u0.free()
    k0.free()
    if args.calibrate_copy:
        out["copy_calibration"] = calibrate_copy(ctx, capi)
        log(f"[bench] copy {out['copy_calibration']}")
    # each leg: the GPU measurement, then (unless --no-cpu) the oracle timed
    # on a bounded sample of the same workload on the host cores
    out["heat64"] = bench_heat64(ctx, capi)
    if not args.no_cpu:
        out["heat64"]["cpu_baseline"] = cpu_heat64()
    log(f"[bench] heat64 {out['heat64']}")
    if not args.no_yee:
        out["yee"] = bench_yee(ctx, capi, **({"nsteps": args.leg_steps} if args.leg_steps else {}))
        if not args.no_cpu:
            out["yee"]["cpu_baseline"] = cpu_yee()
        log(f"[bench] yee {out['yee']}")
    if not args.no_yee_eps:
        out["yee_eps_adjoint"] = bench_yee_eps(ctx, capi, **({"nsteps": args.leg_steps, "snapshots": (4,)}
                                                             if args.leg_steps else {}))
        if not args.no_cpu:
            out["yee_eps_adjoint"]["cpu_baseline"] = cpu_yee_eps()
        log(f"[bench] yee eps {out['yee_eps_adjoint']}")
    if not args.no_pennes:
        out["pennes_adjoint"] = bench_pennes(ctx, capi, **({"nsteps": args.leg_steps} if args.leg_steps else {}))
        if not args.no_cpu:
            out["pennes_adjoint"]["cpu_baseline"] = cpu_pennes()
        log(f"[bench] pennes {out['pennes_adjoint']}")
    if not args.no_batch:
        out["batch2d"] = bench_batch(ctx, capi, **({"nsteps": args.leg_steps} if args.leg_steps else {}))
        if not args.no_cpu:
            from masterthesis_amd.workloads import batch2d_ensemble
            nb = args.leg_steps or 500
            out["batch2d"]["cpu_baseline"] = cpu_batch(batch2d_ensemble(4096, nb, seed=0), nb)
        log(f"[bench] batch {out['batch2d']}")
    if not args.no_emfdtd:
        out["emfdtd"] = bench_emfdtd(ctx, capi, **({"nsteps": args.leg_steps} if args.leg_steps else {}))
        if not args.no_cpu:
            out["emfdtd"]["cpu_baseline"] = cpu_emfdtd()
        log(f"[bench] emfdtd {out['emfdtd']}")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
