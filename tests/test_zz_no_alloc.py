"""scafes_yee_run_noalloc, the entry point that allocates nothing, against
scafes_yee_run (same kernels): fields and FP64 energy records bit for bit,
on two calls in a row, the second on a smaller box (the scratch array keeps
the first call's partials; nothing may read them), in the two-step,
fused single-step and in-place schedules; a box whose energy partials do
not fit the scratch array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from masterthesis_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(hip):
    c = capi.Context(0)
    yield c
    c.close()


def _run_both(ctx, ncell, code, dt, nsteps, every, second_buf):
    nf = (ncell + 1) ** 3
    out = []
    for noalloc in (False, True):
        f = ctx.alloc(6 * nf, dt)
        f2 = ctx.alloc(6 * nf, dt) if second_buf else None
        capi.check(ctx.L.scafes_yee_init_pulse(ctx.h, ncell, code, f.ptr, 3.0))
        nrec = (nsteps + every - 1) // every
        which = C.c_int(-1)
        if noalloc:
            edev = ctx.alloc(nrec, np.float64)
            edev.fill(-1.0)
            capi.check(ctx.L.scafes_yee_run_noalloc(ctx.h, ncell, code, 0.5, f.ptr, f2.ptr if f2 else None, nsteps,
                                                    every, edev.ptr, C.byref(which)))
            ctx.sync()
            e = edev.read()
            edev.free()
        else:
            e = np.zeros(nrec)
            capi.check(ctx.L.scafes_yee_run(ctx.h, ncell, code, 0.5, f.ptr, f2.ptr if f2 else None, nsteps, every,
                                            e.ctypes.data_as(C.POINTER(C.c_double)), C.byref(which)))
        res = (f2 if (f2 is not None and which.value) else f).read()
        out.append((res, e))
        f.free()
        if f2 is not None:
            f2.free()
    return out


@pytest.mark.parametrize("code,dt", [(capi.F32, np.float32), (capi.F64, np.float64)])
@pytest.mark.parametrize("second_buf", [True, False])
def test_noalloc_equals_allocating_twice_then_smaller(ctx, code, dt, second_buf):
    for ncell, nsteps in ((48, 9), (20, 7)):
        (fa, ea), (fb, eb) = _run_both(ctx, ncell, code, dt, nsteps, 2, second_buf)
        np.testing.assert_array_equal(fb, fa)
        np.testing.assert_array_equal(eb, ea)
        assert ea[0] > 0.0 and np.all(np.isfinite(ea))


def test_noalloc_refuses_box_beyond_scratch(ctx):
    small = ctx.alloc(16, np.float32)
    edev = ctx.alloc(4, np.float64)
    which = C.c_int(-1)
    rc = ctx.L.scafes_yee_run_noalloc(ctx.h, 1300, capi.F32, 0.5, small.ptr, edev.ptr, 4, 1, edev.ptr,
                                      C.byref(which))
    assert rc == -1
    assert b"scratch" in ctx.L.scafes_last_error()
    # without energy records no partials are needed, and bad arguments are
    # still refused before any launch
    rc = ctx.L.scafes_yee_run_noalloc(ctx.h, 12, capi.F32, 0.9, small.ptr, None, 1, 0, None, C.byref(which))
    assert rc == -1 and b"Courant" in ctx.L.scafes_last_error()
    small.free()
    edev.free()
