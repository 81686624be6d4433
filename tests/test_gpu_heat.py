"""GPU parity of the heat-family kernels against the CPU oracle, through the C ABI.

FP64 kernels follow the CPU path's operation order and are compared
bit-for-bit (the 1e-12 bar is met with zero difference).  FP32 paths are
compared against the FP64 oracle with the tolerance written in each test.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from masterthesis_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(hip):
    c = capi.Context(0)
    yield c
    c.close()


def _xs(n):
    return [np.array([(1.0 / (m - 1)) * i for i in range(m)]) for m in n]


def timelin_known(n, t, pennes=False, rhob_cb_w=1.0, border=None):
    """F, G, Y of the TimeLinSpaceLin problems (stencil-0 layout, (z,y,x)),
    evaluated with the hooks' operation order; border: None (Dirichlet) or
    one of the flux classes (lambda = alpha = epsilon = sigma = 1)."""
    bc = border
    dim = len(n)
    xs = _xs(n)
    shape = tuple(reversed(n))
    mesh = np.meshgrid(*reversed(xs), indexing="ij")[::-1]  # mesh[p] indexed (z,y,x)
    s = np.ones(shape)
    for p in range(dim):
        s = s + mesh[p]
    border = np.zeros(shape, bool)
    idx = np.meshgrid(*[np.arange(m) for m in reversed(n)], indexing="ij")[::-1]
    for p in range(dim):
        border |= (idx[p] < 1) | (idx[p] > n[p] - 2)
    Y = s * (1.0 + t)
    if bc is None:
        F = np.where(border, 0.0, s * (1.0 * 1.0))
        if pennes:
            F = np.where(border, 0.0, F + rhob_cb_w * (s * (1.0 + t)))
        G = np.where(border, s * (1.0 + t), 0.0)
        return F, G, Y
    F = s * (1.0 * 1.0)
    if pennes:
        F = F + rhob_cb_w * (s * (1.0 + t))
    near = np.zeros(shape, bool)
    for p in range(dim):
        near |= idx[p] == 0
    if bc == capi.BORDER_NEUMANN:
        flux = np.full(shape, -1.0 * 1.0 * (1.0 + t))
    elif bc == capi.BORDER_CAUCHY:
        flux = s * (1.0 * (1.0 + t)) + 1.0 * (1.0 + t)
    else:
        q = s * (1.0 + t)
        flux = (q * q * q * q) * (1.0 * 1.0) + 1.0 * (1.0 + t)
    G = np.where(near, Y, np.where(border, flux, 0.0))
    return F, G, Y


@pytest.mark.parametrize("n,pennes,bc", [((16, 16, 16), False, None), ((12, 10, 9), False, None),
                                         ((20, 17), False, None), ((33,), False, None),
                                         ((14, 13, 12), True, None),
                                         ((14, 13, 12), False, capi.BORDER_NEUMANN),
                                         ((11, 15), False, capi.BORDER_CAUCHY),
                                         ((13, 12, 14), True, capi.BORDER_NEUMANN),
                                         ((12, 14, 13), True, capi.BORDER_CAUCHY),
                                         ((12, 11, 13), True, capi.BORDER_RADIATION),
                                         ((27,), True, capi.BORDER_RADIATION)])
def test_timelin_bitexact(ctx, n, pennes, bc):
    """TimeLinSpaceLin heat/Pennes with Dirichlet or flux borders: FP64 device
    steps equal the oracle bit for bit and stay on the exact solution."""
    dim = len(n)
    g = capi.geom(n)
    og = O.make_grid(n)
    h2 = sum(1.0 / O.lib().oracle_gridsize(og, p) ** 2 for p in range(dim))
    tau = 0.8 / (2.0 * h2 + (1.0 if pennes else 0.0))
    if bc is not None:
        hmin = min(O.lib().oracle_gridsize(og, p) for p in range(dim))
        tau = 0.4 / (2.0 * h2 + 1.0 + (8.0 * (1.0 + dim) ** 3 / hmin if bc == capi.BORDER_RADIATION else 0.0))
    nsteps = 25
    coef = O.heat7_coef(tau, perf_c=1.0, perf_w=1.0, rhob_cb_w=1.0, has_perf=pennes, has_source=True)
    ref, linf, _ = O.heat7_run(og, np.zeros(O.memall_shape(og)), nsteps, 0.0, coef,
                               O.KNOWN_PENNES_TIMELIN if pennes else O.KNOWN_HEAT_TIMELIN,
                               O.BORDER_FIELD if bc is None else bc)
    # initial state: (1+sum x)(1+t0) on nodes, ghosts 0
    _, _, Y0 = timelin_known(n, 0.0)
    u0 = np.zeros(O.memall_shape(og))
    O.nodes_view(og, u0)[...] = Y0.reshape(O.nodes_view(og, u0).shape)
    ne = u0.size
    a, b = ctx.upload(u0), ctx.upload(u0)
    Fd, God, Gd = ctx.alloc(Y0.size), ctx.alloc(Y0.size), ctx.alloc(Y0.size)
    p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, perf_c=1.0, perf_w=1.0,
                         has_perf=int(pennes), has_source=1,
                         border_mode=capi.BORDER_FIELD if bc is None else bc,
                         border_value=0.0, lam=1.0, alpha=1.0, epsilon=1.0, sigma=1.0)
    for step in range(1, nsteps + 1):
        F_old, G_old, _ = timelin_known(n, (step - 1) * tau, pennes, border=bc)
        _, G_new, _ = timelin_known(n, step * tau, pennes, border=bc)
        Fd.write(F_old)
        God.write(G_old)
        Gd.write(G_new)
        capi.check(ctx.L.scafes_heat7_step_f64(ctx.h, capi.C.byref(g), 1, a.ptr, b.ptr, Fd.ptr, God.ptr, Gd.ptr,
                                               capi.C.byref(p)))
        a, b = b, a
    got = a.read().reshape(u0.shape)
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, ref))
    assert got.size == ne
    # and the device path reproduces the exact discrete solution
    _, _, Yn = timelin_known(n, nsteps * tau)
    assert np.abs(O.nodes_view(og, got) - Yn.reshape(O.nodes_view(og, got).shape)).max() < 1e-12


def sine_state(og):
    n = [og.n[p] for p in range(og.dim)]
    u = np.zeros(O.memall_shape(og))
    xs = _xs(n)
    v = np.ones(tuple(reversed(n)))
    mesh = np.meshgrid(*reversed(xs), indexing="ij")[::-1]
    for p in range(og.dim):
        v = v * np.sin(np.pi * mesh[p])
    O.nodes_view(og, u)[...] = v.reshape(O.nodes_view(og, u).shape)
    return u, v


def test_sine_mode_64cube_bitexact_and_analytic(ctx):
    """Heat 64^3, FP64, tau = 0.9 h^2/6, 100 steps, u0 = prod sin(pi x), G = 0."""
    n = 64
    og = O.make_grid((n, n, n))
    g = capi.geom((n, n, n))
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 6.0
    nsteps = 100
    u0, prod = sine_state(og)
    ref, _, _ = O.heat7_run(og, u0, nsteps, 0.0, O.heat7_coef(tau, coeff_a=1.0), O.KNOWN_NONE)
    a, b = ctx.upload(u0), ctx.upload(u0)
    p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, has_perf=0, has_source=0,
                         border_mode=capi.BORDER_VALUE, border_value=0.0)
    for _ in range(nsteps):
        capi.check(ctx.L.scafes_heat7_step_f64(ctx.h, capi.C.byref(g), 1, a.ptr, b.ptr, None, None, None,
                                               capi.C.byref(p)))
        a, b = b, a
    got = a.read().reshape(u0.shape)
    np.testing.assert_array_equal(got, ref)
    # error against the analytic solution exp(-3 pi^2 t) prod sin(pi x) (reported)
    t = nsteps * tau
    exact = math.exp(-3 * math.pi ** 2 * t) * prod
    rel = np.abs(O.nodes_view(og, got) - exact.reshape(O.nodes_view(og, got).shape)).max() / np.abs(exact).max()
    print(f"\n[config heat64] rel. error vs analytic solution: {rel:.3e}")
    assert rel < 1e-3


@pytest.mark.parametrize("nsteps", [0, 1, 37])
def test_heat7_run_f64_equals_steps(ctx, nsteps):
    """scafes_heat7_run_f64: nsteps launches in one call, bit-identical to the
    oracle (and so to nsteps single-step calls), the result in u2 after an
    odd count; update rules that read known fields are refused."""
    n = (33, 30, 28)
    og = O.make_grid(n)
    g = capi.geom(n)
    h = 1.0 / (n[0] - 1)
    tau = 0.9 * h * h / 6.0
    u0, _ = sine_state(og)
    u0 = u0 + 0.01 * np.random.default_rng(nsteps).random(u0.shape)
    coef = O.heat7_coef(tau, coeff_a=1.0)
    C = capi.C
    for bm, bv in ((capi.BORDER_VALUE, 0.25), (capi.BORDER_KEEP, 0.0)):
        ref = u0
        for _ in range(nsteps):
            ref = O.heat7_step(og, ref, border=bm, border_value=bv, coef=coef)
        a, b = ctx.upload(u0), ctx.upload(u0)
        p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, has_perf=0, has_source=0, border_mode=bm,
                             border_value=bv)
        which = C.c_int(-1)
        capi.check(ctx.L.scafes_heat7_run_f64(ctx.h, C.byref(g), 1, a.ptr, b.ptr, nsteps, C.byref(p),
                                              C.byref(which)))
        assert which.value == nsteps % 2
        got = (b if which.value else a).read().reshape(u0.shape)
        np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, ref))
        a.free()
        b.free()
    a, b = ctx.upload(u0), ctx.upload(u0)
    err_arg = -1  # SCAFES_ERR_ARG
    for kw in ({"has_source": 1}, {"border_mode": capi.BORDER_NEUMANN, "lam": 1.0}):
        p = capi.Heat7Params(tau=tau, coeff_a=1.0, rhoc_inv=1.0, **kw)
        assert ctx.L.scafes_heat7_run_f64(ctx.h, C.byref(g), 1, a.ptr, b.ptr, 3, C.byref(p), None) == err_arg
    a.free()
    b.free()


def heatvar_inputs(n, seed=0, amp=1e-3):
    og = O.make_grid(n)
    xs = _xs(n)
    mesh = np.meshgrid(*reversed(xs), indexing="ij")[::-1]
    kap = np.zeros(O.memall_shape(og))
    K = 1.0 + 0.5 * np.sin(2 * np.pi * mesh[0]) * np.sin(2 * np.pi * mesh[1])
    O.nodes_view(og, kap)[...] = K.reshape(O.nodes_view(og, kap).shape)
    u, _ = sine_state(og)
    rng = np.random.default_rng(seed)
    noise = amp * (2 * rng.random(tuple(reversed(n))) - 1)
    inner = np.zeros(tuple(reversed(n)), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    O.nodes_view(og, u)[...] += np.where(inner, noise, 0.0).reshape(O.nodes_view(og, u).shape)
    return og, u, kap


def run_heatvar(ctx, g, mode, u0, kap, tau, nsteps, cfg=-1, dtype=np.float64, border=capi.BORDER_VALUE,
                bval=0.0):
    a, b = ctx.upload(u0.astype(dtype)), ctx.upload(u0.astype(dtype))
    kd = ctx.upload(kap.astype(dtype))
    c1 = c2 = None
    if mode == capi.HEATVAR_F32_KAHAN:
        c1, c2 = ctx.alloc(u0.size, np.float32), ctx.alloc(u0.size, np.float32)
        c1.fill(0.0)
        c2.fill(0.0)
    which = capi.C.c_int(0)
    capi.check(ctx.L.scafes_heat7var_run(ctx.h, capi.C.byref(g), 1, mode, a.ptr, b.ptr,
                                         c1.ptr if c1 else None, c2.ptr if c2 else None, kd.ptr, tau,
                                         nsteps, border, bval, cfg, capi.C.byref(which)))
    res = (b if which.value else a).read().reshape(u0.shape)
    comp = None
    if c1 is not None:
        comp = (c2 if which.value else c1).read().reshape(u0.shape)
    for d in (a, b, kd, c1, c2):
        if d is not None:
            d.free()
    return res, comp


NUM_CFGS = capi.load().scafes_heat7var_num_cfgs()


def cfg_steps_per_sweep(cfg):
    st = capi.C.c_int(0)
    capi.check(capi.load().scafes_heat7var_cfg_desc(cfg, None, None, None, None, capi.C.byref(st)))
    return st.value


@pytest.mark.parametrize("cfg", list(range(NUM_CFGS)))
def test_heatvar_f64_bitexact_all_tiles(ctx, cfg):
    """Every tile configuration, one- and two-steps-per-sweep, even and odd step counts."""
    n = (37, 29, 23)  # not multiples of any tile edge
    og, u0, kap = heatvar_inputs(n, seed=cfg)
    g = capi.geom(n)
    tau = 0.9 / (2.0 * 1.5 * sum((m - 1.0) ** 2 for m in n))  # 1/tau >= 2 kappa_max sum 1/h^2
    nsteps = 12 + cfg % 2
    ref = O.heat7var_run(og, u0, kap, tau, nsteps)
    got, _ = run_heatvar(ctx, g, capi.HEATVAR_F64, u0, kap, tau, nsteps, cfg=cfg)
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, ref))


TB_CFGS = [c for c in range(NUM_CFGS) if cfg_steps_per_sweep(c) >= 2]


@pytest.mark.parametrize("mode", [capi.HEATVAR_F32, capi.HEATVAR_F32_KAHAN, capi.HEATVAR_F64])
@pytest.mark.parametrize("cfg", TB_CFGS)
def test_heatvar_multi_step_sweep_matches_one_step(ctx, cfg, mode):
    """Two or three steps per sweep perform the one-step kernel's arithmetic node by node:
    identical bits in every precision, including the Kahan compensation, with
    nonzero kept borders, several z chunks and a tile grid that does not
    divide the domain."""
    n = (70, 45, 150)
    og, u0, kap = heatvar_inputs(n, seed=11)
    rng = np.random.default_rng(5)
    border = np.ones(tuple(reversed(n)), bool)
    border[1:-1, 1:-1, 1:-1] = False
    O.nodes_view(og, u0)[...] += np.where(border, 0.25 + rng.random(border.shape), 0.0)
    g = capi.geom(n)
    tau = 0.9 / (2.0 * 1.5 * sum((m - 1.0) ** 2 for m in n))
    dt =np.float64 if mode == capi.HEATVAR_F64 else np.float32
    nsteps = 9
    one, c_one = run_heatvar(ctx, g, mode, u0, kap, tau, nsteps, cfg=0, dtype=dt, border=capi.BORDER_KEEP)
    two, c_two = run_heatvar(ctx, g, mode, u0, kap, tau, nsteps, cfg=cfg, dtype=dt, border=capi.BORDER_KEEP)
    np.testing.assert_array_equal(O.nodes_view(og, two), O.nodes_view(og, one))
    if mode == capi.HEATVAR_F32_KAHAN:
        np.testing.assert_array_equal(O.nodes_view(og, c_two), O.nodes_view(og, c_one))
    # border value 0.5 imposed on every step
    one, _ = run_heatvar(ctx, g, mode, u0, kap, tau, 4, cfg=0, dtype=dt, border=capi.BORDER_VALUE, bval=0.5)
    two, _ = run_heatvar(ctx, g, mode, u0, kap, tau, 4, cfg=cfg, dtype=dt, border=capi.BORDER_VALUE, bval=0.5)
    np.testing.assert_array_equal(O.nodes_view(og, two), O.nodes_view(og, one))


def test_heatvar_fp32_kahan_error(ctx):
    """FP32 + Kahan stays within 1e-5 relative (to max|u|) of the FP64 oracle."""
    n = (48, 48, 48)
    og, u0, kap = heatvar_inputs(n, seed=3)
    g = capi.geom(n)
    h = 1.0 / (n[0] - 1)
    tau = 0.9 * h * h / (6 * 1.5)
    nsteps = 300
    ref = O.heat7var_run(og, u0, kap, tau, nsteps)
    k32, comp = run_heatvar(ctx, g, capi.HEATVAR_F32_KAHAN, u0, kap, tau, nsteps, dtype=np.float32)
    p32, _ = run_heatvar(ctx, g, capi.HEATVAR_F32, u0, kap, tau, nsteps, dtype=np.float32)
    scale = np.abs(O.nodes_view(og, ref)).max()
    ek = np.abs(O.nodes_view(og, k32.astype(np.float64) + comp) - O.nodes_view(og, ref)).max() / scale
    ep = np.abs(O.nodes_view(og, p32.astype(np.float64)) - O.nodes_view(og, ref)).max() / scale
    print(f"\n[heat7var 48^3 x {nsteps}] rel err fp32={ep:.2e} fp32+kahan={ek:.2e}")
    assert ek < 1e-5
    assert ek < ep


def test_heatvar_init_matches_numpy(ctx):
    """Device-side synthetic init (sine + splitmix64 noise, kappa field)."""
    n = (20, 18, 16)
    og = O.make_grid(n)
    g = capi.geom(n)
    ne = capi.memall_count(g, 1)
    u, k = ctx.alloc(ne), ctx.alloc(ne)
    u.fill(0.0)
    k.fill(0.0)
    capi.check(ctx.L.scafes_heat7var_init(ctx.h, capi.C.byref(g), 1, capi.F64, u.ptr, k.ptr, 7, 1e-3))
    from masterthesis_amd.workloads import heatvar_init_numpy
    u_np, k_np = heatvar_init_numpy(n, seed=7, amp=1e-3)
    got_u = O.nodes_view(og, u.read().reshape(O.memall_shape(og)))
    got_k = O.nodes_view(og, k.read().reshape(O.memall_shape(og)))
    np.testing.assert_allclose(got_u, u_np, rtol=0, atol=1e-15)
    np.testing.assert_allclose(got_k, k_np, rtol=0, atol=1e-15)


def test_reductions(ctx):
    n = (21, 19, 17)
    og = O.make_grid(n)
    g = capi.geom(n)
    rng = np.random.default_rng(1)
    a = rng.standard_normal(O.memall_shape(og))
    b = rng.standard_normal(tuple(reversed(n)))  # stencil-0 layout
    da, db = ctx.upload(a), ctx.upload(b)
    out = capi.C.c_double(0)
    capi.check(ctx.L.scafes_reduce(ctx.h, capi.C.byref(g), capi.F64, da.ptr, 1, db.ptr, 0,
                                   capi.RED_MAX_ABS_DIFF, 0, capi.C.byref(out)))
    an = O.nodes_view(og, a)
    assert out.value == np.abs(an - b).max()
    capi.check(ctx.L.scafes_reduce(ctx.h, capi.C.byref(g), capi.F64, da.ptr, 1, db.ptr, 0,
                                   capi.RED_SUM_SQ_DIFF, 1, capi.C.byref(out)))
    inner = (slice(1, -1),) * 3
    assert abs(out.value - np.sum((an[inner] - b[inner]) ** 2)) < 1e-10 * out.value


@pytest.mark.parametrize("nbytes", [0, 1, 15, 16, 17, 1023, 4096, 65541, 1 << 20, (64 << 20) + 7])
@pytest.mark.parametrize("offs", [(0, 0), (4, 4), (16, 48), (3, 0)])
def test_device_copy_sizes_and_alignment(ctx, nbytes, offs):
    """scafes_memcpy_d2d byte for byte against numpy: 16-byte path (aligned
    ends, several elements per thread, grids from one block to 65536), the
    byte tail, and unaligned ends (byte path only); bytes outside the target
    range keep their values."""
    so, do = offs
    total = nbytes + 64
    rng = np.random.default_rng(nbytes + 7 * so + do)
    src_h = rng.integers(0, 256, total, dtype=np.uint8)
    dst_h = rng.integers(0, 256, total, dtype=np.uint8)
    src, dst = ctx.upload(src_h), ctx.upload(dst_h)
    vp = capi.C.c_void_p
    capi.check(ctx.L.scafes_memcpy_d2d(ctx.h, vp(dst.ptr.value + do), vp(src.ptr.value + so), nbytes))
    want = dst_h.copy()
    want[do:do + nbytes] = src_h[so:so + nbytes]
    np.testing.assert_array_equal(dst.read(), want)
    src.free()
    dst.free()


# ---------------------------------------------------------------------------
# Pennes field form (forward and adjoint steps)

def pennes_inputs(n, seed=11):
    og = O.make_grid(n)
    h = 1.0 / (n[0] - 1)
    rng = np.random.default_rng(seed)
    w = 0.5 * np.exp(0.3 * rng.standard_normal(O.memall_shape(og)))
    x = [np.linspace(0, 1, m) for m in n]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    q = np.zeros(O.memall_shape(og))
    O.nodes_view(og, q)[...] = 50.0 * np.exp(-((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) / (2 * .15 ** 2))
    tau = 0.9 / (6 * 0.5 / h ** 2 + w.max())
    coef = O.pennes_coef(tau=tau, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=37.0)
    T0 = np.full(O.memall_shape(og), 37.0)
    return og, w, q, coef, T0


def test_pennes_forward_and_adjoint_bitexact(ctx):
    n = (19, 17, 15)
    og, w, q, coef, T0 = pennes_inputs(n)
    g = capi.geom(n)
    N = 15
    st = O.pennes_forward_all(og, T0, w, q, coef, N)
    rng = np.random.default_rng(4)
    target = st[-1] + 0.01 * rng.standard_normal(st[-1].shape)
    J, grad = O.pennes_grad(og, st, w, target, coef, N)
    pp = capi.PennesParams(tau=coef.tau, cdiff=coef.cdiff, cperf=coef.cperf, csrc=coef.csrc, t_a=coef.t_a)
    a, b = ctx.upload(T0), ctx.upload(T0)
    dw, dq = ctx.upload(w), ctx.upload(q)
    for _ in range(N):
        capi.check(ctx.L.scafes_pennes_step(ctx.h, capi.C.byref(g), 1, capi.F64, a.ptr, b.ptr, dw.ptr,
                                            dq.ptr, capi.C.byref(pp)))
        a, b = b, a
    np.testing.assert_array_equal(a.read().reshape(T0.shape), st[-1])
    # adjoint sweep over the oracle's stored states
    inner = np.zeros(T0.shape, bool)
    O.nodes_view(og, inner)[1:-1, 1:-1, 1:-1] = True
    lam0 = np.where(inner, st[-1] - target, 0.0)
    la, lb = ctx.upload(lam0), ctx.upload(lam0)
    dg = ctx.upload(np.zeros_like(T0))
    dprev = ctx.alloc(T0.size)
    for s in range(N, 0, -1):
        dprev.write(st[s - 1])
        capi.check(ctx.L.scafes_pennes_adjoint_step(ctx.h, capi.C.byref(g), 1, capi.F64, la.ptr, lb.ptr,
                                                    dw.ptr, dprev.ptr, dg.ptr, capi.C.byref(pp)))
        la, lb = lb, la
    np.testing.assert_array_equal(dg.read().reshape(T0.shape), grad)


def test_pennes_fp32_deviation_form_error(ctx):
    """FP32 stream in the deviation variable theta = T - T_a (T_a = 0 in the
    kernel) against the FP64 oracle on absolute temperatures."""
    n = (32, 32, 32)
    og, w, q, coef, T0 = pennes_inputs(n)
    g = capi.geom(n)
    N = 200
    st = O.pennes_forward_all(og, T0, w, q, coef, N)
    pp = capi.PennesParams(tau=coef.tau, cdiff=coef.cdiff, cperf=coef.cperf, csrc=coef.csrc, t_a=0.0)
    th0 = (T0 - coef.t_a).astype(np.float32)
    a, b = ctx.upload(th0), ctx.upload(th0)
    dw, dq = ctx.upload(w.astype(np.float32)), ctx.upload(q.astype(np.float32))
    for _ in range(N):
        capi.check(ctx.L.scafes_pennes_step(ctx.h, capi.C.byref(g), 1, capi.F32, a.ptr, b.ptr, dw.ptr,
                                            dq.ptr, capi.C.byref(pp)))
        a, b = b, a
    got = a.read().reshape(T0.shape).astype(np.float64) + coef.t_a
    rise = np.abs(st[-1] - coef.t_a).max()
    err = np.abs(got - st[-1]).max() / rise
    print(f"\n[pennes 32^3 x {N}] fp32 rel err (to max temperature rise) {err:.2e}")
    assert err < 1e-5
