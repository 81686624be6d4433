"""The HIP kernels against the golden fixtures in tests/golden (frozen
outputs of the pinned oracle on small seeded inputs), through the C ABI:
FP64 families bit for bit on the node block, the FP32 batched 2-D ensemble
within the tolerances written below."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from masterthesis_amd import capi

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


def fixture(name):
    return MANIFEST[name]["params"], np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def ctx(hip):
    c = capi.Context(0)
    yield c
    c.close()


def ping_pong(ctx, u0, nsteps, step):
    a, b = ctx.upload(u0), ctx.upload(u0)
    for _ in range(nsteps):
        step(a, b)
        a, b = b, a
    out = a.read().reshape(u0.shape)
    a.free()
    b.free()
    return out


def test_heat7_sine(ctx):
    par, f = fixture("heat7_sine_16")
    n = par["n"]
    g, og = capi.geom((n, n, n)), O.make_grid((n, n, n))
    p = capi.Heat7Params(tau=par["tau"], coeff_a=1.0, rhoc_inv=1.0, border_mode=capi.BORDER_VALUE)
    got = ping_pong(ctx, f["u0"], par["nsteps"], lambda a, b: capi.check(ctx.L.scafes_heat7_step_f64(
        ctx.h, C.byref(g), 1, a.ptr, b.ptr, None, None, None, C.byref(p))))
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, f["u"]))


def test_heat7var_fp64(ctx):
    par, f = fixture("heat7var_16")
    n = par["n"]
    g, og = capi.geom((n, n, n)), O.make_grid((n, n, n))
    a, b, k = ctx.upload(f["u0"]), ctx.upload(f["u0"]), ctx.upload(f["kappa"])
    second = C.c_int(-1)
    capi.check(ctx.L.scafes_heat7var_run(ctx.h, C.byref(g), 1, capi.HEATVAR_F64, a.ptr, b.ptr, None, None, k.ptr,
                                         par["tau"], par["nsteps"], capi.BORDER_KEEP, 0.0, -1, C.byref(second)))
    got = (b if second.value else a).read().reshape(f["u0"].shape)
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, f["u"]))


@pytest.mark.parametrize("snaps", [3, 9])
def test_pennes_forward_and_gradient(ctx, snaps):
    par, f = fixture("pennes_grad_12")
    n, N = par["n"], par["nsteps"]
    g, og = capi.geom((n, n, n)), O.make_grid((n, n, n))
    pp = capi.PennesParams(tau=par["tau"], cdiff=par["cdiff"], cperf=par["cperf"], csrc=par["csrc"], t_a=par["t_a"])
    dw, dq = ctx.upload(f["w"]), ctx.upload(f["q"])
    got = ping_pong(ctx, f["T0"], N, lambda a, b: capi.check(ctx.L.scafes_pennes_step(
        ctx.h, C.byref(g), 1, capi.F64, a.ptr, b.ptr, dw.ptr, dq.ptr, C.byref(pp))))
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, f["T"]))
    dT, dtg, dg = ctx.upload(f["T0"]), ctx.upload(f["target"]), ctx.alloc(f["T0"].size)
    J = C.c_double(0)
    capi.check(ctx.L.scafes_pennes_gradient(ctx.h, C.byref(g), 1, capi.F64, dT.ptr, dw.ptr, dq.ptr, dtg.ptr,
                                            C.byref(pp), N, snaps, dg.ptr, C.byref(J), None))
    inner = np.zeros(f["T0"].shape, bool)
    O.nodes_view(og, inner)[1:-1, 1:-1, 1:-1] = True
    np.testing.assert_array_equal(np.where(inner, dg.read().reshape(f["T0"].shape), 0), f["grad"])
    assert abs(J.value - par["J"]) <= 1e-12 * abs(par["J"])


def test_yee_fp64(ctx):
    par, f = fixture("yee_12")
    N = par["ncell"]
    d, d2 = ctx.upload(f["f0"]), ctx.alloc(f["f0"].size)
    second = C.c_int(-1)
    capi.check(ctx.L.scafes_yee_run(ctx.h, N, capi.F64, par["courant"], d.ptr, d2.ptr, par["nsteps"], 0, None,
                                    C.byref(second)))
    got = (d2 if second.value else d).read().reshape(f["f0"].shape)
    np.testing.assert_array_equal(got, f["f"])


def test_yee_eps_fp64(ctx):
    """Yee with a permittivity field: the forward state and dJ/deps bit for
    bit (the gradient with 3 snapshots), J to 1e-12."""
    par, f = fixture("yee_eps_10")
    N, n = par["ncell"], par["nsteps"]
    d, de, w = ctx.upload(f["f0"]), ctx.upload(f["eps"]), ctx.alloc(f["f0"].size)
    capi.check(ctx.L.scafes_yee_eps_run(ctx.h, N, capi.F64, par["courant"], de.ptr, d.ptr, w.ptr, n))
    np.testing.assert_array_equal(d.read().reshape(f["f0"].shape), f["f"])
    d0, dt, dg = ctx.upload(f["f0"]), ctx.upload(f["target"]), ctx.alloc(f["eps"].size)
    J = C.c_double(0)
    capi.check(ctx.L.scafes_yee_eps_gradient(ctx.h, N, capi.F64, par["courant"], d0.ptr, de.ptr, dt.ptr, n, 3,
                                             dg.ptr, C.byref(J), None))
    np.testing.assert_array_equal(dg.read().reshape(f["eps"].shape), f["grad"])
    assert abs(J.value - par["J"]) <= 1e-12 * abs(par["J"])
    for b in (d, de, w, d0, dt, dg):
        b.free()


def test_heat2d_batch_fp32(ctx):
    """FP32 state against the FP64 fixture: fields within 2e-5 of their
    maximum; J and dJ/dkappa within 1e-3 relative (the misfit is a small
    difference of O(1) values, so FP32 rounding is amplified there)."""
    par, f = fixture("heat2d_batch")
    sizes, N = par["sizes"], par["nsteps"]
    n = np.array(sizes, dtype=np.int32)
    off = np.zeros(len(sizes), dtype=np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    u0 = np.concatenate([f[f"u0_{m}"].ravel() for m in sizes]).astype(np.float32)
    tgt = np.concatenate([f[f"target_{m}"].ravel() for m in sizes]).astype(np.float32)
    order = np.argsort(-((n + 31) // 32), kind="stable").astype(np.int32)
    d = {k: ctx.upload(v) for k, v in (("n", n), ("off", off), ("kappa", f["kappa"].astype(np.float32)),
                                       ("tau", f["tau"].astype(np.float32)), ("order", order), ("u", u0),
                                       ("t", tgt))}
    J, gr = ctx.alloc(len(sizes)), ctx.alloc(len(sizes))
    capi.check(ctx.L.scafes_batch2d_run(ctx.h, len(sizes), d["n"].ptr, d["off"].ptr, d["kappa"].ptr, d["tau"].ptr,
                                        d["order"].ptr, d["u"].ptr, d["t"].ptr, J.ptr, gr.ptr, N, 0))
    out, Jd, gd = d["u"].read(), J.read(), gr.read()
    for i, m in enumerate(sizes):
        ref = f[f"u_{m}"]
        got = out[off[i]:off[i] + m * m].reshape(m, m)
        assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max(), m
        assert abs(Jd[i] - f["J"][i]) <= 1e-3 * abs(f["J"][i]), m
        assert abs(gd[i] - f["dJ_dkappa"][i]) <= 1e-3 * abs(f["dJ_dkappa"][i]), m


def test_tumor_fp64(ctx):
    par, f = fixture("tumor_13x9x8")
    n, ext = tuple(par["n"]), par["extent"]
    og = O.make_grid(n, x0=[0.0] * 3, x1=ext)
    h = [O.lib().oracle_gridsize(C.byref(og), q) for q in range(3)]
    g = capi.geom(n, h=h)
    rc = par["rho"] * par["c"]
    p = capi.PennesParams(tau=par["tau"], cdiff=par["k"] / rc, cperf=(par["rho_b"] * par["c_pb"]) / rc,
                          csrc=1.0 / rc, t_a=par["t_a"])
    p.border_mode, p.k_cond, p.h_conv, p.t_inf, p.q_bc = capi.BORDER_TUMOR, par["k"], par["h"], par["t_inf"], par["q_bc"]
    dw, dq = ctx.upload(f["w"]), ctx.upload(f["q"])
    got = ping_pong(ctx, f["u0"], par["nsteps"], lambda a, b: capi.check(ctx.L.scafes_pennes_step(
        ctx.h, C.byref(g), 1, capi.F64, a.ptr, b.ptr, dw.ptr, dq.ptr, C.byref(p))))
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, f["u"]))


def test_mri_fp64(ctx):
    par, f = fixture("mri_11x9x8")
    n, ext = tuple(par["n"]), par["extent"]
    og = O.make_grid(n, x0=[0.0] * 3, x1=ext)
    h = [O.lib().oracle_gridsize(C.byref(og), q) for q in range(3)]
    g = capi.geom(n, h=h)
    p = capi.MriParams(tau=par["tau"], h_conv=par["h_conv"], t_amb=par["t_amb"], q_bc=par["q_bc"],
                       q_skull=par["q_skull"], epsilon=par["epsilon"], sigma=par["sigma"])
    mats = [ctx.upload(np.ascontiguousarray(f["mat_" + m])) for m in O.MRI_FIELDS]
    ptrs = (capi.vp * 9)(*[m.ptr for m in mats])
    coef = ctx.alloc(capi.MRI_NCOEF * f["u0"].size)
    capi.check(ctx.L.scafes_mri_coef_f64(ctx.h, C.byref(g), 1, ptrs, C.byref(p), coef.ptr))
    got = ping_pong(ctx, f["u0"], par["nsteps"], lambda a, b: capi.check(ctx.L.scafes_mri_step_f64(
        ctx.h, C.byref(g), 1, a.ptr, b.ptr, ptrs, coef.ptr, C.byref(p))))
    np.testing.assert_array_equal(O.nodes_view(og, got), O.nodes_view(og, f["u"]))


def test_emfdtd_fp64(ctx):
    par, f = fixture("emfdtd_16")
    p = capi.emfdtd_params(par["n"], d=par["d"], src=tuple(par["src"]), sink=tuple(par["sink"]))
    dev = ctx.upload(np.ascontiguousarray(f["state0"]))
    sink = ctx.alloc(par["nsteps"])
    capi.check(ctx.L.scafes_emfdtd_run(ctx.h, C.byref(p), dev.ptr, par["nsteps"], 0.0, sink.ptr))
    got = dev.read().reshape(f["state0"].shape)
    np.testing.assert_array_equal(got[:O.EM_STATE], f["state"])
    np.testing.assert_array_equal(sink.read(), f["sink"])
