"""The Tumor example's border rule (Neumann heat flux on every face, Cauchy
convection on the far face of the last axis) in the Pennes kernel, against
the oracle's restatement of TumorHeatEqnFDM::updateInner/updateBorder:
FP64 bit for bit in 1-D, 2-D and 3-D with one and two border layers, FP32
within the tolerance written below; the example class end to end through
Problem::iterateOverTime, including its NetCDF output at the end; and the
perfusion gradient through these borders (adjoint step and checkpointed
gradient FP64 bit for bit against the oracle, FP32 within 1e-4)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle_lib as O
from masterthesis_amd import capi

pytestmark = pytest.mark.gpu
ERR_ARG = -1  # SCAFES_ERR_ARG
BIN = os.path.join(O.ROOT, "masterthesis_amd", "bin")
EXTENT = (0.06, 0.03, 0.03)


@pytest.fixture(scope="module")
def ctx(hip):
    c = capi.Context(0)
    yield c
    c.close()


def _case(n, nlayers=1, q_bc=120.0):
    dim = len(n)
    g = O.make_grid(n, x0=[0.0] * dim, x1=list(EXTENT[:dim]), nlayers=nlayers)
    h = [O.lib().oracle_gridsize(C.byref(g), p) for p in range(dim)]
    rho, cc, K, H = 1040.0, 3650.0, 0.6, 10.0
    tau = 0.5 / (sum(2.0 * K / (rho * cc) / hp ** 2 for hp in h) + 2.0 * H / (rho * cc * h[-1]) +
                 1052.0 * 3800.0 / (rho * cc) * 0.004)
    center = [e / 2.0 for e in EXTENT[:dim - 1]] + [EXTENT[dim - 1] - 0.008]
    coef = O.tumor_coef(tau, center, 0.005, q_bc=q_bc)
    return g, h, tau, coef


def _params(coef, shift=0.0):
    p = capi.PennesParams(tau=coef.tau, cdiff=coef.k / (coef.rho * coef.c),
                          cperf=(coef.rho_b * coef.c_pb) / (coef.rho * coef.c), csrc=1.0 / (coef.rho * coef.c),
                          t_a=coef.t_a - shift)
    p.border_mode, p.k_cond, p.h_conv, p.t_inf, p.q_bc = capi.BORDER_TUMOR, coef.k, coef.h, coef.t_inf - shift, coef.q_bc
    return p


@pytest.mark.parametrize("n", [(31,), (25, 13), (21, 11, 12)])
@pytest.mark.parametrize("nlayers", [1, 2])
def test_tumor_border_fp64_bitexact(ctx, n, nlayers):
    g, h, tau, coef = _case(n, nlayers)
    w, q = O.tumor_fields(g, coef)
    rng = np.random.default_rng(11)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = 37.0 + rng.standard_normal(tuple(reversed(n)))
    geo = capi.geom(n, h=h, nlayers=nlayers)
    p = _params(coef)
    a, b, dw, dq = ctx.upload(u), ctx.upload(u), ctx.upload(w), ctx.upload(q)
    ref = u
    for _ in range(25):
        capi.check(ctx.L.scafes_pennes_step(ctx.h, C.byref(geo), 1, capi.F64, a.ptr, b.ptr, dw.ptr, dq.ptr,
                                            C.byref(p)))
        a, b = b, a
        ref = O.tumor_step(g, ref, coef)
    got = a.read().reshape(ref.shape)
    np.testing.assert_array_equal(O.nodes_view(g, got), O.nodes_view(g, ref))


def test_tumor_border_fp32_deviation_state(ctx):
    """FP32 on theta = T - T_a (the form the host engine uses) within 2e-5 K
    of the FP64 oracle after 100 steps."""
    n = (21, 11, 12)
    g, h, tau, coef = _case(n)
    w, q = O.tumor_fields(g, coef)
    u = np.full(O.memall_shape(g), 37.0)
    ref = u
    for _ in range(100):
        ref = O.tumor_step(g, ref, coef)
    geo = capi.geom(n, h=h)
    p = _params(coef, shift=coef.t_a)
    th = (u - coef.t_a).astype(np.float32)
    a, b = ctx.upload(th), ctx.upload(th)
    dw, dq = ctx.upload(w.astype(np.float32)), ctx.upload(q.astype(np.float32))
    for _ in range(100):
        capi.check(ctx.L.scafes_pennes_step(ctx.h, C.byref(geo), 1, capi.F32, a.ptr, b.ptr, dw.ptr, dq.ptr,
                                            C.byref(p)))
        a, b = b, a
    got = a.read().reshape(ref.shape).astype(np.float64) + coef.t_a
    err = np.abs(O.nodes_view(g, got) - O.nodes_view(g, ref)).max()
    assert err < 2e-5, err
    assert np.abs(O.nodes_view(g, ref) - 37.0).max() > 1e-3  # the state moved


def _ghost_nan(g, inner_values):
    a = np.full(O.memall_shape(g), np.nan)
    O.nodes_view(g, a)[...] = inner_values
    return a


@pytest.mark.parametrize("n", [(31,), (25, 13), (21, 11, 12)])
@pytest.mark.parametrize("nlayers", [1, 2])
def test_tumor_adjoint_step_fp64_bitexact(ctx, n, nlayers):
    """One adjoint step under the Tumor border: lam' on every node equals the
    oracle's transposed step (pinned against the assembled matrix) bit for
    bit, ghost layers NaN so that a read beyond the grid would show; the
    gradient term lam*(tau*cperf*(T_a - T_prev)) lands on every node."""
    g, h, tau, coef = _case(n, nlayers)
    w, q = O.tumor_fields(g, coef)
    p = _params(coef)
    oc = O.pennes_coef(tau=p.tau, cdiff=p.cdiff, cperf=p.cperf, csrc=p.csrc, t_a=p.t_a)
    ob = O.tumor_border(coef.k, coef.h, coef.t_inf, coef.q_bc)
    rng = np.random.default_rng(23)
    shape = tuple(reversed(n))
    lam = _ghost_nan(g, rng.standard_normal(shape))
    tp = _ghost_nan(g, 37.0 + rng.standard_normal(shape))
    g0 = np.zeros(O.memall_shape(g))
    O.nodes_view(g, g0)[...] = rng.standard_normal(shape)
    geo = capi.geom(n, h=h, nlayers=nlayers)
    dl, dn, dw, dt, dg = ctx.upload(lam), ctx.upload(np.zeros_like(lam)), ctx.upload(w), ctx.upload(tp), \
        ctx.upload(g0)
    capi.check(ctx.L.scafes_pennes_adjoint_step(ctx.h, C.byref(geo), 1, capi.F64, dl.ptr, dn.ptr, dw.ptr, dt.ptr,
                                                dg.ptr, C.byref(p)))
    want = O.pennes_tumor_adjoint_step(g, lam, w, oc, ob)
    got = dn.read().reshape(lam.shape)
    np.testing.assert_array_equal(O.nodes_view(g, got), O.nodes_view(g, want))
    gwant = O.nodes_view(g, g0) + O.nodes_view(g, lam) * ((p.tau * p.cperf) * (p.t_a - O.nodes_view(g, tp)))
    np.testing.assert_array_equal(O.nodes_view(g, dg.read().reshape(lam.shape)), gwant)


def _grad_setup(n, nsteps, nlayers=1):
    g, h, tau, coef = _case(n, nlayers)
    w, q = O.tumor_fields(g, coef)
    p = _params(coef)
    oc = O.pennes_coef(tau=p.tau, cdiff=p.cdiff, cperf=p.cperf, csrc=p.csrc, t_a=p.t_a)
    ob = O.tumor_border(coef.k, coef.h, coef.t_inf, coef.q_bc)
    rng = np.random.default_rng(31)
    t0 = np.zeros(O.memall_shape(g))
    O.nodes_view(g, t0)[...] = 37.0 + 0.5 * rng.standard_normal(tuple(reversed(n)))
    w_t = w * (1.0 + 0.3 * rng.standard_normal(w.shape))
    target = O.pennes_tumor_forward_all(g, t0, w_t, q, oc, ob, nsteps)[-1]
    return g, h, coef, w, q, oc, ob, t0, target


def _device_grad(ctx, n, h, nlayers, dtype, coef, t0, w, q, target, nsteps, snaps, shift=0.0):
    geo = capi.geom(n, h=h, nlayers=nlayers)
    dt = np.float64 if dtype == capi.F64 else np.float32
    dT, dw, dq, dtg = (ctx.upload((a - s).astype(dt)) for a, s in ((t0, shift), (w, 0), (q, 0), (target, shift)))
    dg = ctx.alloc(t0.size, np.float64)
    J = C.c_double(0)
    stats = np.zeros(4)
    capi.check(ctx.L.scafes_pennes_gradient(ctx.h, C.byref(geo), 1, dtype, dT.ptr, dw.ptr, dq.ptr, dtg.ptr,
                                            C.byref(_params(coef, shift)), nsteps, snaps, dg.ptr, C.byref(J),
                                            stats.ctypes.data_as(C.POINTER(C.c_double))))
    grad = dg.read().reshape(t0.shape)
    for d in (dT, dw, dq, dtg, dg):
        d.free()
    return J.value, grad, stats


@pytest.mark.parametrize("n,nlayers", [((12, 10, 9), 1), ((15, 11), 1), ((31,), 1), ((13, 9, 8), 2)])
@pytest.mark.parametrize("snaps", [0, 2, 5, 29])
def test_tumor_gradient_fp64_bitexact(ctx, n, nlayers, snaps):
    """dJ/dw through the Tumor border: every checkpoint budget reproduces the
    oracle's store-all adjoint (itself pinned by finite differences) bit for
    bit on every node, ghost layers zero."""
    N = 30
    g, h, coef, w, q, oc, ob, t0, target = _grad_setup(n, N, nlayers)
    st = O.pennes_tumor_forward_all(g, t0, w, q, oc, ob, N)
    Jref, gref = O.pennes_tumor_grad(g, st, w, target, oc, ob, N)
    J, grad, stats = _device_grad(ctx, n, h, nlayers, capi.F64, coef, t0, w, q, target, N, snaps)
    np.testing.assert_array_equal(grad, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert np.abs(O.nodes_view(g, grad)[0]).max() > 0  # the border nodes carry sensitivity
    assert stats[0] == ctx.L.scafes_checkpoint_forward_steps(N, snaps) and stats[1] == N


@pytest.mark.parametrize("N,snaps", [(9, 8), (11, 10), (12, 3)])
def test_tumor_multi_step_reverse_sweeps_ragged(ctx, N, snaps):
    """The multi-step reverse sweeps under the Tumor border on a grid with
    ragged x and y tiles and two z chunks: store-all (four steps per sweep,
    then the remainder) and a binomial budget, FP64 bit for bit against the
    oracle on every node."""
    n = (70, 33, 40)
    g, h, coef, w, q, oc, ob, t0, target = _grad_setup(n, N)
    st = O.pennes_tumor_forward_all(g, t0, w, q, oc, ob, N)
    Jref, gref = O.pennes_tumor_grad(g, st, w, target, oc, ob, N)
    J, grad, stats = _device_grad(ctx, n, h, 1, capi.F64, coef, t0, w, q, target, N, snaps)
    np.testing.assert_array_equal(grad, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert stats[1] == N


@pytest.mark.parametrize("dtype", [capi.F32, capi.F32_KAHAN])
def test_tumor_gradient_fp32_deviation_state(ctx, dtype):
    """FP32 fields on theta = T - T_a, FP64 gradient: within 1e-4 (relative
    L2) of the FP64 oracle, J within 1e-4."""
    n, N = (21, 11, 12), 40
    g, h, coef, w, q, oc, ob, t0, target = _grad_setup(n, N)
    st = O.pennes_tumor_forward_all(g, t0, w, q, oc, ob, N)
    Jref, gref = O.pennes_tumor_grad(g, st, w, target, oc, ob, N)
    J, grad, _ = _device_grad(ctx, n, h, 1, dtype, coef, t0, w, q, target, N, 7, shift=coef.t_a)
    err = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    assert err < 1e-4, err
    assert J == pytest.approx(Jref, rel=1e-4)


def test_tumor_adjoint_rejects_zero_conductivity(ctx):
    """k_cond = 0 would divide the flux terms by zero: the adjoint step and
    the gradient refuse it with SCAFES_ERR_ARG before any launch."""
    n = (12, 10, 9)
    g, h, tau, coef = _case(n)
    geo = capi.geom(n, h=h)
    p = _params(coef)
    p.k_cond = 0.0
    ne = int(np.prod(O.memall_shape(g)))
    bufs = [ctx.alloc(ne, np.float64) for _ in range(5)]
    J = C.c_double(0)
    rc = ctx.L.scafes_pennes_adjoint_step(ctx.h, C.byref(geo), 1, capi.F64, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                          bufs[3].ptr, bufs[4].ptr, C.byref(p))
    assert rc == ERR_ARG and b"k_cond" in ctx.L.scafes_last_error()
    rc = ctx.L.scafes_pennes_gradient(ctx.h, C.byref(geo), 1, capi.F64, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                      bufs[3].ptr, C.byref(p), 3, 2, bufs[4].ptr, C.byref(J), None)
    assert rc == ERR_ARG and b"k_cond" in ctx.L.scafes_last_error()


def test_tumor_autograd_vs_finite_differences(ctx):
    """PennesMisfit with the Tumor borders (FP64): J.backward() against
    central differences of the device forward along random directions over
    every node (error ~ eps^2: 4.6e-9 relative at eps = 1e-5 in the oracle)."""
    torch = pytest.importorskip("torch")
    from masterthesis_amd.autograd import PennesMisfit
    n, N = (21, 11, 12), 30
    g, h, coef, w, q, oc, ob, t0, target = _grad_setup(n, N)
    fn = PennesMisfit(ctx, n, oc.tau, oc.cdiff, oc.cperf, oc.csrc, oc.t_a, N, snapshots=5, dtype="f64",
                      tumor=(coef.k, coef.h, coef.t_inf, coef.q_bc), h=h)
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    J = fn(wt, torch.tensor(t0), torch.tensor(q), torch.tensor(target))
    J.backward()
    gw = wt.grad.numpy()
    dom = np.zeros(w.shape, bool)
    O.nodes_view(g, dom)[...] = True
    rng = np.random.default_rng(6)
    for trial in range(2):
        v = np.where(dom, w * rng.standard_normal(w.shape), 0.0)
        eps = 1e-5
        with torch.no_grad():
            jp = fn(torch.tensor(w + eps * v), torch.tensor(t0), torch.tensor(q), torch.tensor(target)).item()
            jm = fn(torch.tensor(w - eps * v), torch.tensor(t0), torch.tensor(q), torch.tensor(target)).item()
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gw * v))
        print(f"\n[tumor FD] trial {trial}: adjoint {ad:.10e} fd {fd:.10e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)


TUMOR_INI = """[Parameters]
RHO = 1040.0
C = 3650.0
K = 0.6
RHO_B = 1052.0
C_PB = 3800.0
H = 10.0
T_INF = 22.4
Q_BC = {q_bc}
DIAMETER = 0.01
DEPTH = 0.008
T_A = 37.0
T_I = 37.0
OMEGA_B_BRAIN = 0.004
OMEGA_B_TUMOR = 0.0007
Q_BRAIN = 25000.0
Q_TUMOR = 25000.0
"""


@pytest.mark.parametrize("n", [(31,), (25, 13), (21, 11, 12)])
def test_tumor_example_through_problem(tmp_path, hip, n):
    dim = len(n)
    nsteps = 40
    g, h, tau, coef = _case(n, q_bc=60.0)
    ini = tmp_path / "Parameters.ini"
    ini.write_text(TUMOR_INI.format(q_bc=60.0))
    out = str(tmp_path / "tumor")
    dump = str(tmp_path / "T.bin")
    args = [f"--dim={dim}", "--nNodes=" + "x".join(map(str, n)), "--coordNodeFirst=" + "x".join(["0"] * dim),
            "--coordNodeLast=" + "x".join(repr(e) for e in EXTENT[:dim]), f"--nTimesteps={nsteps}",
            f"--tau={tau!r}", f"--configfile={ini}", f"--outputfile={out}", "--threshold=0"]
    env = dict(os.environ, SCAFES_DUMP=dump)
    p = subprocess.run([os.path.join(BIN, f"tumor_{dim}d")] + args, capture_output=True, text=True, env=env,
                       timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert f"LAST_ITER {nsteps}" in p.stdout
    ref = np.full(O.memall_shape(g), 37.0)
    for _ in range(nsteps):
        ref = O.tumor_step(g, ref, coef)
    np.testing.assert_array_equal(np.fromfile(dump, dtype=np.float64), O.nodes_view(g, ref).ravel())
    nc = netcdf_file(out + ".nc", "r", mmap=False)
    T = nc.variables["TNewDom"]
    assert T.shape[0] == 1  # AT_END: one record
    np.testing.assert_array_equal(T[0].ravel(), O.nodes_view(g, ref).ravel())
    nc.close()
