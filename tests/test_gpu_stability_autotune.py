"""Explicit-step stability checks and the heat7var autotuner on the GPU.

Stability: the diffusion entry points refuse a tau above the positivity
bound the reference's drivers check (tests/system/HeatEqnFDM/Dirichlet/
TimeLinSpaceLin/main3D.cpp:52-55; examples/TumorHeatEqnFDM/
startSimulation.py:238-301), 1/tau >= sum_p 2 d_max / h_p^2 + reaction,
and accept tau just below it.  The coefficient maxima are device
reductions kept per pointer: a field rewritten through the library is
re-read.

Autotuner: it returns a configuration that runs on the grid, keeps it for
that grid only, and the tuned run is bit-identical to the default one
(every configuration computes the same arithmetic).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from masterthesis_amd import capi

pytestmark = pytest.mark.gpu
ERR_ARG = -1


@pytest.fixture(scope="module")
def ctx(hip):
    c = capi.Context(0)
    yield c
    c.close()


def _rc_msg(ctx, rc):
    return rc, ctx.L.scafes_last_error().decode()


def test_heat7_rejects_unstable_tau(ctx):
    n = (12, 10, 9)
    g = capi.geom(n)
    bound = 1.0 / sum(2.0 / g.h[p] ** 2 for p in range(3))
    ne = capi.memall_count(g, 1)
    a, b = ctx.upload(np.zeros(ne)), ctx.upload(np.zeros(ne))
    for tau, ok in ((0.999 * bound, True), (1.001 * bound, False)):
        p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, border_mode=capi.BORDER_VALUE)
        rc, msg = _rc_msg(ctx, ctx.L.scafes_heat7_step_f64(ctx.h, C.byref(g), 1, a.ptr, b.ptr, None, None, None,
                                                           C.byref(p)))
        assert (rc == 0) == ok, msg
        if not ok:
            assert rc == ERR_ARG and "unstable" in msg
    # the perfusion term counts: stable without it, unstable with it
    p = capi.Heat7Params(tau=0.95 * bound, coeff_a=1.0, rhoc_inv=1.0, perf_c=1.0, perf_w=0.2 / bound, has_perf=1,
                         border_mode=capi.BORDER_VALUE)
    rc, msg = _rc_msg(ctx, ctx.L.scafes_heat7_step_f64(ctx.h, C.byref(g), 1, a.ptr, b.ptr, None, None, None,
                                                       C.byref(p)))
    assert rc == ERR_ARG and "unstable" in msg


@pytest.mark.parametrize("mode", [capi.HEATVAR_F64, capi.HEATVAR_F32, capi.HEATVAR_F32_KAHAN])
def test_heat7var_rejects_unstable_tau_and_sees_rewrites(ctx, mode):
    n = (20, 18, 16)
    g = capi.geom(n)
    ne = capi.memall_count(g, 1)
    dt = np.float64 if mode == capi.HEATVAR_F64 else np.float32
    og = O.make_grid(n)
    kap = np.ones(O.memall_shape(og))
    O.nodes_view(og, kap)[...] = 0.5 + np.random.default_rng(3).random(O.nodes_view(og, kap).shape)
    kmax = float(np.abs(O.nodes_view(og, kap)).max())
    bound = 1.0 / sum(2.0 * kmax / g.h[p] ** 2 for p in range(3))
    u, u2, k = ctx.upload(np.zeros(ne, dt)), ctx.upload(np.zeros(ne, dt)), ctx.upload(kap.ravel().astype(dt))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.upload(np.zeros(ne, np.float32)), ctx.upload(np.zeros(ne, np.float32))
    which = C.c_int(0)

    def run(tau, nsteps=4):
        rc = ctx.L.scafes_heat7var_run(ctx.h, C.byref(g), 1, mode, u.ptr, u2.ptr, c1.ptr if c1 else None,
                                       c2.ptr if c2 else None, k.ptr, tau, nsteps, capi.BORDER_VALUE, 0.0, -1,
                                       C.byref(which))
        return _rc_msg(ctx, rc)

    rc, msg = run(0.99 * bound)
    assert rc == 0, msg
    rc, msg = run(1.01 * bound)
    assert rc == ERR_ARG and "unstable" in msg, msg
    # kappa doubled through the library: the cached maximum is dropped
    k.write((2.0 * kap.ravel()).astype(dt))
    rc, msg = run(0.99 * bound)
    assert rc == ERR_ARG and "unstable" in msg, msg
    rc, msg = run(0.49 * bound)
    assert rc == 0, msg
    # no steps, no check
    rc, msg = run(10.0 * bound, nsteps=0)
    assert rc == 0, msg


@pytest.mark.parametrize("border", [capi.BORDER_KEEP, capi.BORDER_TUMOR])
def test_pennes_rejects_unstable_tau(ctx, border):
    n = (14, 12, 10)
    g = capi.geom(n)
    ne = capi.memall_count(g, 1)
    og = O.make_grid(n)
    w = np.zeros(O.memall_shape(og))
    O.nodes_view(og, w)[...] = 0.2 + np.random.default_rng(5).random(O.nodes_view(og, w).shape)
    wmax = float(O.nodes_view(og, w).max())
    cdiff, cperf, K, H = 0.5, 2.0, 0.6, 10.0
    rate = sum(2.0 * cdiff / g.h[p] ** 2 for p in range(3)) + cperf * wmax
    if border == capi.BORDER_TUMOR:
        rate += 2.0 * cdiff * H / (K * g.h[2])
    a, b = ctx.upload(np.zeros(ne)), ctx.upload(np.zeros(ne))
    wd, qd = ctx.upload(w.ravel()), ctx.upload(np.zeros(ne))
    for tau, ok in ((0.995 / rate, True), (1.005 / rate, False)):
        p = capi.PennesParams(tau=tau, cdiff=cdiff, cperf=cperf, csrc=1.0, t_a=37.0, border_mode=border,
                              k_cond=K, h_conv=H, t_inf=20.0, q_bc=1.0)
        rc, msg = _rc_msg(ctx, ctx.L.scafes_pennes_step(ctx.h, C.byref(g), 1, capi.F64, a.ptr, b.ptr, wd.ptr, qd.ptr,
                                                        C.byref(p)))
        assert (rc == 0) == ok, msg
        if not ok:
            assert rc == ERR_ARG and "unstable" in msg


def test_gradients_check_tau(ctx):
    """The gradient entry points check the bound on every call (their
    coefficient fields change between the calls of an optimisation)."""
    n = (10, 9, 8)
    g = capi.geom(n)
    ne = capi.memall_count(g, 1)
    og = O.make_grid(n)
    kap = np.ones(O.memall_shape(og))
    bound = 1.0 / sum(2.0 / g.h[p] ** 2 for p in range(3))
    u0, k, tgt = ctx.upload(np.zeros(ne)), ctx.upload(kap.ravel()), ctx.upload(np.zeros(ne))
    grad = ctx.alloc(ne)
    J = C.c_double(0)
    for tau, ok in ((0.9 * bound, True), (1.1 * bound, False)):
        rc = ctx.L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, capi.HEATVAR_F64, u0.ptr, k.ptr, tgt.ptr, tau,
                                            capi.BORDER_KEEP, 0.0, 3, 1, grad.ptr, C.byref(J), None)
        assert (rc == 0) == ok, ctx.L.scafes_last_error()
    w = ctx.upload(np.full(ne, 1.0))
    q = ctx.upload(np.zeros(ne))
    rate = sum(2.0 * 0.5 / g.h[p] ** 2 for p in range(3)) + 1.0
    for tau, ok in ((0.9 / rate, True), (1.1 / rate, False)):
        p = capi.PennesParams(tau=tau, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=37.0)
        rc = ctx.L.scafes_pennes_gradient(ctx.h, C.byref(g), 1, capi.F64, u0.ptr, w.ptr, q.ptr, tgt.ptr, C.byref(p), 3,
                                          1, grad.ptr, C.byref(J), None)
        assert (rc == 0) == ok, ctx.L.scafes_last_error()
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))


def _heatvar_case(n, seed):
    og = O.make_grid(n)
    rng = np.random.default_rng(seed)
    u0 = np.zeros(O.memall_shape(og))
    O.nodes_view(og, u0)[...] = rng.random(O.nodes_view(og, u0).shape)
    kap = np.ones(O.memall_shape(og))
    O.nodes_view(og, kap)[...] = 0.5 + rng.random(O.nodes_view(og, kap).shape)
    tau = 0.9 / (2.0 * 1.5 * sum((m - 1.0) ** 2 for m in n))
    return og, u0, kap, tau


def _run(ctx, g, mode, u0, kap, tau, nsteps, cfg=-1):
    dt = np.float64 if mode == capi.HEATVAR_F64 else np.float32
    a, b, k = ctx.upload(u0.ravel().astype(dt)), ctx.upload(u0.ravel().astype(dt)), ctx.upload(kap.ravel().astype(dt))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.upload(np.zeros(u0.size, np.float32)), ctx.upload(np.zeros(u0.size, np.float32))
    which = C.c_int(0)
    capi.check(ctx.L.scafes_heat7var_run(ctx.h, C.byref(g), 1, mode, a.ptr, b.ptr, c1.ptr if c1 else None,
                                         c2.ptr if c2 else None, k.ptr, tau, nsteps, capi.BORDER_KEEP, 0.0, cfg,
                                         C.byref(which)))
    u = (b if which.value else a).read()
    c = (c2 if which.value else c1).read() if c1 is not None else None
    for d in (a, b, k, c1, c2):
        if d is not None:
            d.free()
    return u, c


@pytest.mark.parametrize("mode", [capi.HEATVAR_F64, capi.HEATVAR_F32_KAHAN])
def test_autotune_picks_valid_cfg_kept_per_grid(ctx, mode):
    n = (70, 66, 40)
    g = capi.geom(n)
    og, u0, kap, tau = _heatvar_case(n, 11)
    ref, cref = _run(ctx, g, mode, u0, kap, tau, 14)
    default = C.c_int(-1)
    capi.check(ctx.L.scafes_heat7var_active_cfg(ctx.h, C.byref(g), 1, mode, C.byref(default)))
    dt = np.float64 if mode == capi.HEATVAR_F64 else np.float32
    ne = u0.size
    a, b, k = ctx.upload(u0.ravel().astype(dt)), ctx.upload(u0.ravel().astype(dt)), ctx.upload(kap.ravel().astype(dt))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.upload(np.zeros(ne, np.float32)), ctx.upload(np.zeros(ne, np.float32))
    best, ms = C.c_int(-1), C.c_float(0)
    capi.check(ctx.L.scafes_autotune_heat7var(ctx.h, C.byref(g), 1, mode, a.ptr, b.ptr, c1.ptr if c1 else None,
                                              c2.ptr if c2 else None, k.ptr, tau, 6, C.byref(best), C.byref(ms)))
    assert 0 <= best.value < ctx.L.scafes_heat7var_num_cfgs() and ms.value > 0.0
    active = C.c_int(-1)
    capi.check(ctx.L.scafes_heat7var_active_cfg(ctx.h, C.byref(g), 1, mode, C.byref(active)))
    assert active.value == best.value
    # another grid keeps its default
    g2 = capi.geom((n[0] + 2, n[1], n[2]))
    other = C.c_int(-1)
    capi.check(ctx.L.scafes_heat7var_active_cfg(ctx.h, C.byref(g2), 1, mode, C.byref(other)))
    assert other.value == default.value
    # the tuned run (cfg = -1 picks it) is bit-identical to the default tile
    got, cgot = _run(ctx, g, mode, u0, kap, tau, 14)
    np.testing.assert_array_equal(got, ref)
    if cref is not None:
        np.testing.assert_array_equal(cgot, cref)
    for d in (a, b, k, c1, c2):
        if d is not None:
            d.free()


@pytest.mark.parametrize("nsteps", [1, 2, 4, 5, 7, 8, 10, 11, 20])
def test_default_sweep_plan_bit_identical(ctx, nsteps):
    """The default run splits the steps into three-step sweeps, a two-step
    sweep for two left over, and (FP32 + Kahan) four-step sweeps for step
    counts that are not a multiple of three: the result equals one-step
    launches bit for bit in every mode."""
    n = (40, 36, 30)
    g = capi.geom(n)
    og, u0, kap, tau = _heatvar_case(n, 21)
    for mode in (capi.HEATVAR_F32_KAHAN, capi.HEATVAR_F32, capi.HEATVAR_F64):
        multi, cm = _run(ctx, g, mode, u0, kap, tau, nsteps)
        one, c1 = _run(ctx, g, mode, u0, kap, tau, nsteps, cfg=0)
        np.testing.assert_array_equal(multi, one)
        if cm is not None:
            np.testing.assert_array_equal(cm, c1)


def _plan(ctx, g, mode, nsteps, cap=64):
    cnt = C.c_int(0)
    buf = (C.c_int * cap)()
    capi.check(ctx.L.scafes_heat7var_plan(ctx.h, C.byref(g), 1, mode, nsteps, buf, cap, C.byref(cnt)))
    return list(buf[:min(cnt.value, cap)]), cnt.value


def test_sweep_plan(ctx):
    """scafes_heat7var_plan reports the default run's sweeps: they add up to
    the step count; FP32 + Kahan covers every count from 6 on with three-
    and four-step sweeps (at most two four-step ones); FP32 and FP64 take a
    remainder of two as one two-step sweep and of one as one step; a 2-D
    grid has no multi-step tile."""
    g = capi.geom((40, 36, 30))
    K, F32, F64 = capi.HEATVAR_F32_KAHAN, capi.HEATVAR_F32, capi.HEATVAR_F64
    for n in range(0, 40):
        p, c = _plan(ctx, g, K, n)
        assert sum(p) == n and c == len(p)
        if n >= 6:
            assert set(p) <= {3, 4} and p.count(4) <= 2, (n, p)
        for mode in (F32, F64):
            q, _ = _plan(ctx, g, mode, n)
            assert sum(q) == n and q[: n // 3] == [3] * (n // 3), (n, q)
            assert q[n // 3:] == ([2] if n % 3 == 2 else [1] if n % 3 == 1 else []), (n, q)
    assert _plan(ctx, g, K, 20)[0] == [4, 4, 3, 3, 3, 3]
    assert _plan(ctx, g, K, 1000, cap=400)[0] == [4] + [3] * 332
    assert _plan(ctx, g, K, 5)[0] == [3, 2]
    assert _plan(ctx, g, K, 4)[0] == [4]
    assert _plan(ctx, capi.geom((40, 36)), K, 5)[0] == [1] * 5
    # the count is reported in full when the array is shorter
    p, c = _plan(ctx, g, K, 20, cap=2)
    assert p == [4, 4] and c == 6
