"""GPU adjoint of the variable-kappa heat w.r.t. the kappa field
(scafes_heat7var_gradient, J = 0.5 sum_inner (u^N - target)^2) against the
oracle's discrete adjoint (pinned by finite differences in
test_oracle_heatvar_grad.py):

 * FP64: gradient bit-identical to the oracle for every snapshot budget, both
   border modes, one and two border layers, 1-D to 3-D; J to 1e-12; on 3-D
   grids the store-all stretches run two reverse steps per sweep, also on
   ragged tiles and several z chunks;
 * FP32 and FP32+Kahan fields (FP64 gradient): within 1e-4 relative L2 of
   the FP64 oracle (measured values printed);
 * FP32: checkpointing changes the schedule, not the arithmetic: the
   gradient with 3 snapshots and with store-all are bit-identical;
 * argument errors;
 * the torch.autograd.Function wrapper (KappaMisfit) against central
   finite differences.
"""
import ctypes as C

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


def _case(n, nlayers=1, seed=0):
    g = O.make_grid(n, nlayers=nlayers)
    rng = np.random.default_rng(seed)
    shape = O.memall_shape(g)
    vs = O.nodes_view(g, np.zeros(shape)).shape
    u0 = np.zeros(shape)
    O.nodes_view(g, u0)[...] = rng.random(vs)
    kap = np.ones(shape)
    O.nodes_view(g, kap)[...] = 0.5 + rng.random(vs)
    h = [1.0 / (m - 1) for m in n]
    tau = 0.8 / sum(2.0 * 1.5 / hp ** 2 for hp in h)
    target = np.zeros(shape)
    O.nodes_view(g, target)[...] = 0.5 * rng.random(vs)
    return g, h, u0, kap, tau, target


def _device(ctx, n, h, nlayers, mode, u0, kap, target, tau, border, bv, N, snaps):
    dt = np.float64 if mode == capi.HEATVAR_F64 else np.float32
    g = capi.geom(n, h=h, nlayers=nlayers)
    du, dk, dtg = ctx.upload(u0.astype(dt)), ctx.upload(kap.astype(dt)), ctx.upload(target.astype(dt))
    dg = ctx.alloc(u0.size)
    J = C.c_double(0)
    stats = np.zeros(4)
    capi.check(ctx.L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, mode, du.ptr, dk.ptr, dtg.ptr, tau, border, bv,
                                              N, snaps, dg.ptr, C.byref(J),
                                              stats.ctypes.data_as(C.POINTER(C.c_double))))
    out = dg.read().reshape(u0.shape)
    for d in (du, dk, dtg, dg):
        d.free()
    return J.value, out, stats


@pytest.mark.parametrize("n,nlayers,border", [((12, 11, 10), 1, capi.BORDER_KEEP), ((12, 11, 10), 1, capi.BORDER_VALUE),
                                              ((13, 12, 11), 2, capi.BORDER_KEEP), ((17, 14), 1, capi.BORDER_VALUE),
                                              ((23,), 1, capi.BORDER_KEEP)])
@pytest.mark.parametrize("snaps", [0, 3, 8])
def test_fp64_bitexact_vs_oracle(ctx, n, nlayers, border, snaps):
    N = 9
    g, h, u0, kap, tau, target = _case(n, nlayers, seed=len(n) + snaps)
    st = O.heat7var_forward_all(g, u0, kap, tau, N, border, 0.3)
    Jref, gref = O.heat7var_grad(g, st, kap, target, tau, N)
    J, grad, stats = _device(ctx, n, h, nlayers, capi.HEATVAR_F64, u0, kap, target, tau, border, 0.3, N, snaps)
    np.testing.assert_array_equal(grad, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert stats[0] == ctx.L.scafes_checkpoint_forward_steps(N, snaps) and stats[1] == N
    assert np.abs(gref).max() > 1e-6


@pytest.mark.parametrize("N,snaps", [(12, 8), (12, 6), (20, 12)])
def test_fp64_peel_runs_bitexact(ctx, N, snaps):
    """Budgets whose binomial schedule splits one step off the left again and
    again (runs of 6, 3 and 8 peeled states, formed three per sweep where
    they can be): FP64 bit for bit against the oracle's store-all adjoint."""
    n = (14, 13, 12)
    g, h, u0, kap, tau, target = _case(n, 1, seed=N + snaps)
    st = O.heat7var_forward_all(g, u0, kap, tau, N, O.BORDER_VALUE, 0.0)
    Jref, gref = O.heat7var_grad(g, st, kap, target, tau, N)
    J, grad, stats = _device(ctx, n, h, 1, capi.HEATVAR_F64, u0, kap, target, tau, capi.BORDER_VALUE, 0.0, N, snaps)
    np.testing.assert_array_equal(grad, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert stats[0] == ctx.L.scafes_checkpoint_forward_steps(N, snaps) and stats[1] == N


@pytest.mark.parametrize("snaps", [0, 8])
def test_fp64_two_step_reverse_sweep_ragged_tiles(ctx, snaps):
    """A 3-D grid with ragged x and y tiles and two z chunks of the
    two-steps-per-sweep adjoint (store-all: every reverse step pair in one
    sweep; no snapshots: one-step launches only): both bit-identical to the
    oracle."""
    n, N = (70, 33, 40), 9
    g, h, u0, kap, tau, target = _case(n, seed=40 + snaps)
    st = O.heat7var_forward_all(g, u0, kap, tau, N, O.BORDER_VALUE, 0.0)
    Jref, gref = O.heat7var_grad(g, st, kap, target, tau, N)
    J, grad, _ = _device(ctx, n, h, 1, capi.HEATVAR_F64, u0, kap, target, tau, capi.BORDER_VALUE, 0.0, N, snaps)
    np.testing.assert_array_equal(grad, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)


@pytest.mark.parametrize("mode", [capi.HEATVAR_F32, capi.HEATVAR_F32_KAHAN])
def test_fp32_modes_vs_fp64_oracle(ctx, mode):
    n, N = (24, 22, 20), 40
    g, h, u0, kap, tau, target = _case(n, seed=9)
    st = O.heat7var_forward_all(g, u0, kap, tau, N, O.BORDER_KEEP, 0.0)
    Jref, gref = O.heat7var_grad(g, st, kap, target, tau, N)
    J, grad, _ = _device(ctx, n, h, 1, mode, u0, kap, target, tau, capi.BORDER_KEEP, 0.0, N, 7)
    rel = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    print(f"\n[var-kappa adjoint {n} x {N}, mode {mode}] gradient rel L2 err {rel:.2e}, "
          f"J rel err {abs(J - Jref) / Jref:.2e}")
    assert rel < 1e-4
    assert abs(J - Jref) < 1e-4 * Jref


@pytest.mark.parametrize("N", [12, 11, 10, 13, 8, 20])
def test_fp32_checkpointing_does_not_change_gradient(ctx, N):
    """FP32 + Kahan: the gradient is the same bit for bit whatever the
    schedule: one-step launches only (no snapshots), a binomial budget (at
    N = 12 with 8 snapshots a run of 6 peeled states, formed three per
    sweep; at N = 20 with 12 a run of 8, formed by two four-step sweeps),
    and store-all, whose reverse steps run 4+4+4, 4+4+3 and 4+4+2 per sweep
    at N = 12, 11, 10 and whose forward states (T^N included) come from
    three- and four-step sweeps: 3+3+3+3 at N = 12, 4+4+3 at 11, 4+3+3 at
    10, 4+3+3+3 at 13, 4+4 at 8 (T^N from a four-step sweep)."""
    n = (40, 36, 33)
    g, h, u0, kap, tau, target = _case(n, seed=4)
    J0, g0, _ = _device(ctx, n, h, 1, capi.HEATVAR_F32_KAHAN, u0, kap, target, tau, capi.BORDER_VALUE, 0.0, N, 0)
    J1, g1, _ = _device(ctx, n, h, 1, capi.HEATVAR_F32_KAHAN, u0, kap, target, tau, capi.BORDER_VALUE, 0.0, N,
                        {12: 8, 20: 12}.get(N, 3))
    J2, g2, _ = _device(ctx, n, h, 1, capi.HEATVAR_F32_KAHAN, u0, kap, target, tau, capi.BORDER_VALUE, 0.0, N, N - 1)
    np.testing.assert_array_equal(g1, g0)
    np.testing.assert_array_equal(g2, g0)
    assert J0 == J1 == J2 > 0


def test_argument_errors(ctx):
    g = capi.geom((8, 8, 8))
    ne = 10 ** 3
    a = ctx.alloc(ne)
    J = C.c_double(0)
    L = ctx.L
    assert L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, 3, a.ptr, a.ptr, a.ptr, 1e-3, 0, 0.0, 2, 1, a.ptr,
                                      C.byref(J), None) != 0
    assert b"mode" in L.scafes_last_error()
    assert L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, 0, a.ptr, a.ptr, a.ptr, 1e-3, capi.BORDER_NEUMANN, 0.0,
                                      2, 1, a.ptr, C.byref(J), None) != 0
    assert L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, 0, a.ptr, a.ptr, a.ptr, 1e-3, 0, 0.0, 0, 1, a.ptr,
                                      C.byref(J), None) != 0
    assert L.scafes_heat7var_gradient(ctx.h, C.byref(g), 1, 0, None, a.ptr, a.ptr, 1e-3, 0, 0.0, 2, 1, a.ptr,
                                      C.byref(J), None) != 0


def test_autograd_function_vs_finite_differences(ctx):
    """KappaMisfit (torch.autograd.Function): J.backward() gives the adjoint
    gradient; a central difference of J along random directions agrees to
    1e-6 (FP64) and to 1e-3 with FP32+Kahan fields."""
    torch = pytest.importorskip("torch")
    from masterthesis_amd.autograd import KappaMisfit
    n, N = (20, 18, 16), 25
    g, h, u0, kap, tau, target = _case(n, seed=12)
    for dtype, tol, eps in (("f64", 1e-6, 1e-5), ("f32_kahan", 1e-3, 1e-2)):
        fn = KappaMisfit(ctx, n, tau, N, snapshots=5, dtype=dtype)
        kt = torch.tensor(kap, dtype=torch.float64, requires_grad=True)
        J = fn(kt, torch.tensor(u0), torch.tensor(target))
        J.backward()
        gk = kt.grad.numpy()
        rng = np.random.default_rng(2)
        nodes = np.zeros(kap.shape, bool)
        O.nodes_view(g, nodes)[...] = True
        v = np.where(nodes, rng.standard_normal(kap.shape), 0.0)
        with torch.no_grad():
            jp = fn(torch.tensor(kap + eps * v), torch.tensor(u0), torch.tensor(target)).item()
            jm = fn(torch.tensor(kap - eps * v), torch.tensor(u0), torch.tensor(target)).item()
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gk * v))
        print(f"\n[kappa autograd {dtype}] adjoint {ad:.10e} fd {fd:.10e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= tol * abs(ad)
