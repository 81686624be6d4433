"""GPU parity of the Yee leapfrog (PEC box, vacuum) against the CPU oracle.

FP64 device path: bit-identical fields, energies equal to round-off.
FP32 device path (the workload's precision): field error and energy drift
against the FP64 oracle within the tolerances written in each test.
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


def random_pec_state(N, seed):
    rng = np.random.default_rng(seed)
    f = O.yee_alloc(N)
    f[0, 1:N, 1:N, :N] = rng.standard_normal((N - 1, N - 1, N))
    f[1, 1:N, :N, 1:N] = rng.standard_normal((N - 1, N, N - 1))
    f[2, :N, 1:N, 1:N] = rng.standard_normal((N, N - 1, N - 1))
    # H on its valid entries too
    f[3, :N, :N, :] = 0.1 * rng.standard_normal((N, N, N + 1))
    f[4, :N, :, :N] = 0.1 * rng.standard_normal((N, N + 1, N))
    f[5, :, :N, :N] = 0.1 * rng.standard_normal((N + 1, N, N))
    return f


def run_device(ctx, N, s, f0, nsteps, dtype, energy_every=0, fused=True):
    """fused: one out-of-place sweep per step (ping-pong buffers); else the
    two in-place half-step sweeps."""
    d = ctx.upload(f0.astype(dtype))
    d2 = ctx.alloc(d.count, dtype) if fused else None
    if d2 is not None:
        d2.fill(-7.0)  # every value must be written by the sweep
    code = capi.F64 if dtype == np.float64 else capi.F32
    nrec = (nsteps + energy_every - 1) // energy_every if energy_every else 0
    e = np.zeros(max(nrec, 1))
    second = capi.C.c_int(-1)
    capi.check(ctx.L.scafes_yee_run(ctx.h, N, code, s, d.ptr, d2.ptr if d2 else None, nsteps, energy_every,
                                    e.ctypes.data_as(capi.C.POINTER(capi.C.c_double)) if nrec else None,
                                    capi.C.byref(second)))
    assert second.value in ((0, 1) if fused else (0,))
    out = (d2 if second.value else d).read().reshape(f0.shape)
    d.free()
    if d2 is not None:
        d2.free()
    return out, e[:nrec]


def oracle_run(N, s, f0, nsteps, energy_every=0):
    f = f0.copy()
    energies = []
    for st in range(nsteps):
        hprev = f[3:].copy()
        O.yee_half_h(N, s, f)
        if energy_every and st % energy_every == 0:
            energies.append(O.yee_energy(N, f, hprev))
        O.yee_half_e(N, s, f)
    return f, np.array(energies)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("N", [16, 37])
def test_yee_fp64_bitexact_with_energy(ctx, N, fused):
    s = 0.5
    f0 = random_pec_state(N, N)
    ref, eref = oracle_run(N, s, f0, 25, energy_every=4)
    got, e = run_device(ctx, N, s, f0, 25, np.float64, energy_every=4, fused=fused)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_allclose(e, eref, rtol=1e-12)
    # the scheme conserves it
    assert np.abs(e - e[0]).max() < 1e-11 * e[0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nsteps", [3, 4])
def test_yee_fused_matches_two_pass(ctx, dtype, nsteps):
    """The fused sweep does the two half steps' arithmetic value by value:
    identical bits, over several z chunks and a tile grid that does not
    divide the box; energies agree to FP64 summation-order round-off."""
    N, s = 150, 0.55
    f0 = random_pec_state(N, 9)
    a, ea = run_device(ctx, N, s, f0, nsteps, dtype, energy_every=1, fused=True)
    b, eb = run_device(ctx, N, s, f0, nsteps, dtype, energy_every=1, fused=False)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(ea, eb, rtol=1e-13)


def test_yee_fp32_error_and_energy_drift(ctx):
    N, s, nsteps = 40, 0.5, 300
    f0 = random_pec_state(N, 3)
    ref, eref = oracle_run(N, s, f0, nsteps, energy_every=50)
    got, e = run_device(ctx, N, s, f0, nsteps, np.float32, energy_every=50)
    err = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
    drift = np.abs(e - eref[0]).max() / eref[0]
    print(f"\n[yee 40^3 x {nsteps}] fp32 rel field err {err:.2e}, energy drift {drift:.2e}")
    assert err < 2e-5
    assert drift < 1e-5


def test_yee_cavity_mode_on_device(ctx):
    """TM mode uniform in z: amplitude follows the exact Yee dispersion."""
    N, s, p, q, nsteps = 24, 0.5, 2, 3, 60
    f = O.yee_alloc(N)
    idx = np.arange(N + 1)
    f[2, :N, :, :] = np.outer(np.sin(np.pi * q * idx / N), np.sin(np.pi * p * idx / N))[None]
    mode = f[2].copy()
    got, _ = run_device(ctx, N, s, f, nsteps, np.float64)
    th = 2 * math.asin(s * math.sqrt(math.sin(math.pi * p / (2 * N)) ** 2 + math.sin(math.pi * q / (2 * N)) ** 2))
    a = math.cos(nsteps * th) - math.tan(th / 2) * math.sin(nsteps * th)
    np.testing.assert_allclose(got[2], a * mode, atol=1e-12)


def test_yee_init_pulse(ctx):
    N, sigma = 20, 3.0
    d = ctx.alloc(6 * (N + 1) ** 3, np.float32)
    capi.check(ctx.L.scafes_yee_init_pulse(ctx.h, N, capi.F32, d.ptr, sigma))
    f = d.read().reshape(6, N + 1, N + 1, N + 1)
    c = N / 2
    k, j, i = np.meshgrid(np.arange(N + 1), np.arange(N + 1), np.arange(N + 1), indexing="ij")
    ez = np.exp(-((i - c) ** 2 + (j - c) ** 2 + (k + 0.5 - c) ** 2) / (2 * sigma ** 2))
    valid = (k < N) & (i >= 1) & (i < N) & (j >= 1) & (j < N)
    np.testing.assert_allclose(f[2], np.where(valid, ez, 0.0), rtol=1e-6, atol=1e-7)
    assert np.abs(f[[0, 1, 3, 4, 5]]).max() == 0.0


def test_yee_rejects_unstable_courant(ctx):
    d = ctx.alloc(6 * 9 ** 3, np.float32)
    with pytest.raises(capi.ScafesError):
        capi.check(ctx.L.scafes_yee_run(ctx.h, 8, capi.F32, 0.6, d.ptr, None, 1, 0, None, None))
    with pytest.raises(capi.ScafesError):  # the second buffer must be distinct
        capi.check(ctx.L.scafes_yee_run(ctx.h, 8, capi.F32, 0.5, d.ptr, d.ptr, 1, 0, None, None))
