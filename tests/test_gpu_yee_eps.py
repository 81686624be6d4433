"""GPU Yee leapfrog with a relative permittivity field eps and its adjoint
gradient dJ/deps (scafes_yee_eps_run, scafes_yee_eps_gradient) against the
oracle (pinned in test_oracle_yee_eps.py):

 * forward FP64: bit-identical to the oracle; eps = 1: bit-identical to the
   vacuum scafes_yee_run in FP64 and FP32; FP32 within 1e-5 relative L2 of
   the FP64 oracle;
 * gradient FP64: bit-identical to the oracle for every snapshot budget and
   box sizes that leave ragged tiles, down to N = 2 and one step; J to 1e-12;
 * the gradient workspace the context keeps: reuse and release do not change
   the bits;
 * gradient FP32 (FP32 fields, FP64 gradient): within 1e-4 relative L2 of
   the FP64 oracle (measured value printed); the snapshot budget changes
   the schedule, not the bits;
 * argument errors;
 * the torch.autograd.Function wrapper (PermittivityMisfit) against a
   central finite difference of the FP64 misfit.
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


def _case(N, seed=0):
    rng = np.random.default_rng(seed)
    m = N + 1
    f0 = rng.standard_normal((6, m, m, m))
    f0[:3] *= O.yee_emask(N)
    eps = 1.0 + 3.0 * rng.random((m, m, m))
    target = 0.3 * rng.standard_normal((3, m, m, m))
    return f0, eps, target


def _run(ctx, N, s, eps, f0, nsteps, dt):
    d, de, w = ctx.upload(f0.astype(dt)), ctx.upload(eps.astype(dt)), ctx.alloc(f0.size, dt)
    w.fill(-7.0)
    code = capi.F64 if dt == np.float64 else capi.F32
    capi.check(ctx.L.scafes_yee_eps_run(ctx.h, N, code, s, de.ptr, d.ptr, w.ptr, nsteps))
    out = d.read().reshape(f0.shape)
    for b in (d, de, w):
        b.free()
    return out


def _grad(ctx, N, s, eps, f0, target, nsteps, snaps, dt):
    bufs = [ctx.upload(a.astype(dt)) for a in (f0, eps, target)]
    g = ctx.alloc(eps.size)
    J = C.c_double(0)
    stats = np.zeros(4)
    code = capi.F64 if dt == np.float64 else capi.F32
    capi.check(ctx.L.scafes_yee_eps_gradient(ctx.h, N, code, s, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, nsteps, snaps,
                                             g.ptr, C.byref(J), stats.ctypes.data_as(C.POINTER(C.c_double))))
    out = g.read().reshape(eps.shape)
    for b in bufs + [g]:
        b.free()
    return J.value, out, stats


@pytest.mark.parametrize("N,nsteps", [(12, 9), (70, 4), (5, 1)])
def test_forward_fp64_bitexact(ctx, N, nsteps):
    s = 0.5
    f0, eps, _ = _case(N, seed=N)
    ref = O.yee_eps_step(N, s, eps, f0.copy(), nsteps)
    np.testing.assert_array_equal(_run(ctx, N, s, eps, f0, nsteps, np.float64), ref)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_unit_eps_is_vacuum(ctx, dt):
    N, s, nsteps = 33, 0.55, 6
    f0, _, _ = _case(N, seed=3)
    f0 = f0.astype(dt)
    d, d2 = ctx.upload(f0), ctx.alloc(f0.size, dt)
    second = C.c_int(-1)
    capi.check(ctx.L.scafes_yee_run(ctx.h, N, capi.F64 if dt == np.float64 else capi.F32, s, d.ptr, d2.ptr, nsteps,
                                    0, None, C.byref(second)))
    vac = (d2 if second.value else d).read().reshape(f0.shape)
    d.free()
    d2.free()
    out = _run(ctx, N, s, np.ones((N + 1,) * 3), f0, nsteps, dt)
    np.testing.assert_array_equal(out, vac)


def test_forward_fp32_vs_fp64_oracle(ctx):
    N, s, nsteps = 40, 0.5, 30
    f0, eps, _ = _case(N, seed=5)
    ref = O.yee_eps_step(N, s, eps, f0.copy(), nsteps)
    out = _run(ctx, N, s, eps, f0, nsteps, np.float32)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print(f"\n[yee eps fp32 {N}^3 x {nsteps}] rel L2 err {rel:.2e}")
    assert rel < 1e-5


@pytest.mark.parametrize("N", [9, 16, 67])
@pytest.mark.parametrize("snaps", [0, 3, 8])
def test_gradient_fp64_bitexact(ctx, N, snaps):
    s, nsteps = 0.5, 9
    f0, eps, target = _case(N, seed=N + snaps)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    Jref, gref = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    J, g, stats = _grad(ctx, N, s, eps, f0, target, nsteps, snaps, np.float64)
    np.testing.assert_array_equal(g, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert stats[0] == ctx.L.scafes_checkpoint_forward_steps(nsteps, snaps) and stats[1] == nsteps
    assert np.abs(gref).max() > 1e-6


@pytest.mark.parametrize("N,nsteps,snaps", [(2, 3, 1), (3, 1, 0), (64, 2, 50), (63, 5, 1)])
def test_gradient_edge_cases_bitexact(ctx, N, nsteps, snaps):
    """Smallest boxes (N = 2: one updated node per E component), one step,
    a snapshot budget above nsteps - 1, and tiles that end one node past a
    64-wide (N = 64) or exactly on one (N = 63)."""
    s = 0.5
    f0, eps, target = _case(N, seed=100 + N)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    Jref, gref = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    J, g, stats = _grad(ctx, N, s, eps, f0, target, nsteps, snaps, np.float64)
    np.testing.assert_array_equal(g, gref)
    assert abs(J - Jref) <= 1e-12 * abs(Jref)
    assert stats[2] == min(snaps, nsteps - 1)
    assert np.count_nonzero(gref) > 0
    np.testing.assert_array_equal(_run(ctx, N, s, eps, f0, nsteps, np.float64), st[nsteps])


def test_gradient_fp32_vs_fp64_oracle(ctx):
    N, s, nsteps = 36, 0.5, 40
    f0, eps, target = _case(N, seed=11)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    Jref, gref = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    J, g, _ = _grad(ctx, N, s, eps, f0, target, nsteps, 6, np.float32)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"\n[yee eps adjoint fp32 {N}^3 x {nsteps}] gradient rel L2 err {rel:.2e}, "
          f"J rel err {abs(J - Jref) / Jref:.2e}")
    assert rel < 1e-4
    assert abs(J - Jref) < 1e-5 * Jref


def test_gradient_fp32_checkpointing_does_not_change_bits(ctx):
    N, s, nsteps = 50, 0.5, 13
    f0, eps, target = _case(N, seed=12)
    J1, g1, _ = _grad(ctx, N, s, eps, f0, target, nsteps, 2, np.float32)
    J2, g2, _ = _grad(ctx, N, s, eps, f0, target, nsteps, nsteps - 1, np.float32)
    np.testing.assert_array_equal(g1, g2)
    assert J1 == J2 > 0


def test_workspace_kept_and_released(ctx):
    """The context keeps the gradient workspace between calls; a smaller
    call reuses it, a release frees it, and the results do not depend on
    either."""
    N, s, nsteps = 24, 0.5, 11
    f0, eps, target = _case(N, seed=31)
    J1, g1, _ = _grad(ctx, N, s, eps, f0, target, nsteps, 10, np.float64)
    J2, g2, _ = _grad(ctx, N, s, eps, f0, target, nsteps, 2, np.float64)
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))
    J3, g3, _ = _grad(ctx, N, s, eps, f0, target, nsteps, 2, np.float64)
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(g2, g3)
    assert J1 == J2 == J3
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))
    capi.check(ctx.L.scafes_ctx_release_workspace(ctx.h))


def test_argument_errors(ctx):
    N = 8
    a = ctx.alloc(6 * (N + 1) ** 3)
    J = C.c_double(0)
    L = ctx.L
    assert L.scafes_yee_eps_gradient(ctx.h, N, capi.F32_KAHAN, 0.5, a.ptr, a.ptr, a.ptr, 2, 1, a.ptr, C.byref(J),
                                     None) != 0
    assert b"dtype" in L.scafes_last_error()
    assert L.scafes_yee_eps_gradient(ctx.h, N, capi.F64, 0.6, a.ptr, a.ptr, a.ptr, 2, 1, a.ptr, C.byref(J),
                                     None) != 0
    assert b"Courant" in L.scafes_last_error()
    assert L.scafes_yee_eps_gradient(ctx.h, 1, capi.F64, 0.5, a.ptr, a.ptr, a.ptr, 2, 1, a.ptr, C.byref(J),
                                     None) != 0
    assert L.scafes_yee_eps_gradient(ctx.h, N, capi.F64, 0.5, a.ptr, a.ptr, a.ptr, 0, 1, a.ptr, C.byref(J),
                                     None) != 0
    assert L.scafes_yee_eps_gradient(ctx.h, N, capi.F64, 0.5, None, a.ptr, a.ptr, 2, 1, a.ptr, C.byref(J),
                                     None) != 0
    assert L.scafes_yee_eps_run(ctx.h, N, capi.F64, 0.5, a.ptr, a.ptr, a.ptr, 2) != 0
    assert b"work" in L.scafes_last_error()
    a.free()


def test_autograd_function_vs_finite_differences(ctx):
    """PermittivityMisfit: J.backward() gives the adjoint gradient.  A central
    difference (h = 1e-5) of the FP64 J along a random direction agrees with
    the FP64 adjoint to 1e-6 and with the FP32-field adjoint to 1e-4.  (A
    difference of the FP32 J itself is no check: its rounding noise over a
    step small enough for the truncation error is of the same size.)"""
    torch = pytest.importorskip("torch")
    from masterthesis_amd.autograd import PermittivityMisfit
    N, s, nsteps, h = 20, 0.5, 25, 1e-5
    f0, eps, target = _case(N, seed=21)
    v = np.random.default_rng(2).standard_normal(eps.shape)
    f64 = PermittivityMisfit(ctx, N, s, nsteps, snapshots=5, dtype="f64")
    with torch.no_grad():
        jp = f64(torch.tensor(eps + h * v), torch.tensor(f0), torch.tensor(target)).item()
        jm = f64(torch.tensor(eps - h * v), torch.tensor(f0), torch.tensor(target)).item()
    fd = (jp - jm) / (2 * h)
    for dtype, tol in (("f64", 1e-6), ("f32", 1e-4)):
        fn = PermittivityMisfit(ctx, N, s, nsteps, snapshots=5, dtype=dtype)
        et = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
        J = fn(et, torch.tensor(f0), torch.tensor(target))
        J.backward()
        ad = float(np.sum(et.grad.numpy() * v))
        print(f"\n[eps autograd {dtype}] adjoint {ad:.10e} fd {fd:.10e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= tol * abs(ad)
