"""ctypes binding of the HIP C ABI (include/scafes_hip.h).

This is the Python side of the boundary used by the test harness and the
benchmark.  It loads masterthesis_amd/lib/libscafes_hip.so and fails loudly
when the library is missing: there is no CPU fallback anywhere in the
package.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libscafes_hip.so")

F32, F64, F32_KAHAN = 0, 1, 2
BORDER_KEEP, BORDER_FIELD, BORDER_VALUE = 0, 1, 2
BORDER_NEUMANN, BORDER_CAUCHY, BORDER_RADIATION = 3, 4, 5
BORDER_TUMOR = 6
HEATVAR_F64, HEATVAR_F32, HEATVAR_F32_KAHAN = 0, 1, 2
RED_MAX_ABS_DIFF, RED_SUM_SQ_DIFF, RED_SUM_PROD = 0, 1, 2


class ScafesError(RuntimeError):
    pass


class Geom(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("nlayers", C.c_int), ("h", C.c_double * 3)]


class Heat7Params(C.Structure):
    _fields_ = [("tau", C.c_double), ("coeff_a", C.c_double), ("rhoc_inv", C.c_double),
                ("perf_c", C.c_double), ("perf_w", C.c_double), ("has_perf", C.c_int),
                ("has_source", C.c_int), ("border_mode", C.c_int), ("border_value", C.c_double),
                ("lam", C.c_double), ("alpha", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double)]


class TimelinParams(C.Structure):
    _fields_ = [("form", C.c_int), ("border_mode", C.c_int), ("x0", C.c_double * 3), ("rho_c", C.c_double),
                ("rhob_cb_w", C.c_double), ("lam", C.c_double), ("alpha", C.c_double), ("epsilon", C.c_double),
                ("sigma", C.c_double)]


KNOWN_HEAT_TIMELIN, KNOWN_PENNES_TIMELIN = 1, 2

EMFDTD_MAX_D = 64
EM_NFIELDS = 26


class EmfdtdParams(C.Structure):
    _fields_ = [("n", C.c_int), ("d", C.c_int), ("cfln", C.c_double), ("dt", C.c_double),
                ("b", C.c_double * (EMFDTD_MAX_D + 1)), ("c", C.c_double * (EMFDTD_MAX_D + 1)),
                ("ook", C.c_double * (EMFDTD_MAX_D + 1)), ("src", C.c_int * 3), ("sink", C.c_int * 3),
                ("f_mod", C.c_double), ("t0", C.c_double), ("spread", C.c_double), ("flags", C.c_int)]


EMFDTD_PSI_SETTLED, EMFDTD_VACUUM_OK = 1, 2


class PennesParams(C.Structure):
    _fields_ = [("tau", C.c_double), ("cdiff", C.c_double), ("cperf", C.c_double),
                ("csrc", C.c_double), ("t_a", C.c_double), ("border_mode", C.c_int),
                ("k_cond", C.c_double), ("h_conv", C.c_double), ("t_inf", C.c_double), ("q_bc", C.c_double)]


class MriParams(C.Structure):
    _fields_ = [("tau", C.c_double), ("h_conv", C.c_double), ("t_amb", C.c_double), ("q_bc", C.c_double),
                ("q_skull", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double)]


MRI_NFIELDS, MRI_NCOEF = 9, 4

_lib = None
vp = C.c_void_p


def _sig(L, name, args, res=C.c_int):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ScafesError(
            f"HIP library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    ip, dp, gp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(Geom)
    _sig(L, "scafes_last_error", [], C.c_char_p)
    _sig(L, "scafes_device_count", [ip])
    _sig(L, "scafes_ctx_create", [C.c_int, C.POINTER(vp)])
    _sig(L, "scafes_ctx_destroy", [vp])
    _sig(L, "scafes_ctx_release_workspace", [vp])
    _sig(L, "scafes_ctx_sync", [vp])
    _sig(L, "scafes_ctx_info", [vp, ip, C.c_char_p, C.c_int])
    _sig(L, "scafes_ctx_stream", [vp, C.POINTER(vp)])
    _sig(L, "scafes_event_record", [vp, C.c_int])
    _sig(L, "scafes_event_elapsed", [vp, C.c_int, C.c_int, C.POINTER(C.c_float)])
    _sig(L, "scafes_malloc", [vp, C.c_size_t, C.POINTER(vp)])
    _sig(L, "scafes_free", [vp, vp])
    _sig(L, "scafes_memcpy_h2d", [vp, vp, vp, C.c_size_t])
    _sig(L, "scafes_memcpy_d2h", [vp, vp, vp, C.c_size_t])
    _sig(L, "scafes_memcpy_d2d", [vp, vp, vp, C.c_size_t])
    _sig(L, "scafes_fill", [vp, vp, C.c_int, C.c_size_t, C.c_double])
    _sig(L, "scafes_convert", [vp, vp, C.c_int, vp, C.c_int, C.c_size_t])
    _sig(L, "scafes_heat7_step_f64", [vp, gp, C.c_int, vp, vp, vp, vp, vp, C.POINTER(Heat7Params)])
    _sig(L, "scafes_heat7var_run", [vp, gp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_double,
                                    C.c_int, C.c_int, C.c_double, C.c_int, ip])
    _sig(L, "scafes_heat7var_num_cfgs", [])
    _sig(L, "scafes_heat7var_cfg_desc", [C.c_int, ip, ip, ip, ip, ip])
    _sig(L, "scafes_heat7var_active_cfg", [vp, gp, C.c_int, C.c_int, ip])
    _sig(L, "scafes_autotune_heat7var", [vp, gp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_double,
                                         C.c_int, ip, C.POINTER(C.c_float)])
    _sig(L, "scafes_heat7var_init", [vp, gp, C.c_int, C.c_int, vp, vp, C.c_uint64, C.c_double])
    _sig(L, "scafes_reduce", [vp, gp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, dp])
    _sig(L, "scafes_coef_changed", [vp, vp])
    _sig(L, "scafes_known_timelin_f64", [vp, gp, C.POINTER(TimelinParams), C.c_double, vp, vp, vp, C.c_int])
    _sig(L, "scafes_pennes_step", [vp, gp, C.c_int, C.c_int, vp, vp, vp, vp, C.POINTER(PennesParams)])
    _sig(L, "scafes_pennes_adjoint_step", [vp, gp, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                           C.POINTER(PennesParams)])
    _sig(L, "scafes_pennes_step_kahan", [vp, gp, C.c_int, vp, vp, vp, vp, vp, vp, C.POINTER(PennesParams)])
    if hasattr(L, "scafes_mri_step_f64"):
        _sig(L, "scafes_mri_coef_f64", [vp, gp, C.c_int, C.POINTER(vp), C.POINTER(MriParams), vp])
        _sig(L, "scafes_mri_step_f64", [vp, gp, C.c_int, vp, vp, C.POINTER(vp), vp, C.POINTER(MriParams)])
    if hasattr(L, "scafes_pennes_gradient"):
        _sig(L, "scafes_pennes_gradient", [vp, gp, C.c_int, C.c_int, vp, vp, vp, vp,
                                           C.POINTER(PennesParams), C.c_int, C.c_int, vp, dp, dp])
        _sig(L, "scafes_checkpoint_forward_steps", [C.c_int, C.c_int], C.c_long)
    if hasattr(L, "scafes_heat7var_gradient"):
        _sig(L, "scafes_heat7var_gradient", [vp, gp, C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_int,
                                             C.c_double, C.c_int, C.c_int, vp, dp, dp])
    if hasattr(L, "scafes_yee_run"):
        _sig(L, "scafes_yee_run", [vp, C.c_int, C.c_int, C.c_double, vp, vp, C.c_int, C.c_int, dp, ip])
        _sig(L, "scafes_yee_run_noalloc", [vp, C.c_int, C.c_int, C.c_double, vp, vp, C.c_int, C.c_int, vp, ip])
        _sig(L, "scafes_yee_init_pulse", [vp, C.c_int, C.c_int, vp, C.c_double])
    if hasattr(L, "scafes_yee_eps_run"):
        _sig(L, "scafes_yee_eps_run", [vp, C.c_int, C.c_int, C.c_double, vp, vp, vp, C.c_int])
        _sig(L, "scafes_yee_eps_gradient", [vp, C.c_int, C.c_int, C.c_double, vp, vp, vp, C.c_int, C.c_int, vp,
                                            dp, dp])
    if hasattr(L, "scafes_emfdtd_run"):
        _sig(L, "scafes_emfdtd_tables", [C.c_int] + [C.c_double] * 6 + [dp, dp, dp])
        _sig(L, "scafes_emfdtd_dt", [C.c_double], C.c_double)
        _sig(L, "scafes_emfdtd_ga_gb", [C.c_size_t, dp, C.c_double, dp, dp])
        _sig(L, "scafes_emfdtd_source", [C.c_double] * 5, C.c_double)
        _sig(L, "scafes_emfdtd_run", [vp, C.POINTER(EmfdtdParams), vp, C.c_int, C.c_double, vp])
        _sig(L, "scafes_emfdtd_vacuum_ok", [vp, C.POINTER(EmfdtdParams), vp, ip])
    if hasattr(L, "scafes_batch2d_run"):
        _sig(L, "scafes_batch2d_run", [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       C.c_int, C.c_int])
        _sig(L, "scafes_batch2d_lds_limit", [ip])
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise ScafesError(f"scafes error {rc}: {load().scafes_last_error().decode()}")


def device_count():
    n = C.c_int(0)
    rc = load().scafes_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def geom(n, h=None, nlayers=1):
    n = list(n)
    g = Geom()
    g.dim = len(n)
    for p in range(3):
        g.n[p] = n[p] if p < len(n) else 1
        if h is None:
            g.h[p] = 1.0 / (n[p] - 1) if p < len(n) else 1.0
        else:
            g.h[p] = h[p] if p < len(n) else 1.0
    g.nlayers = nlayers
    return g


def memall_count(g, hw):
    c = 1
    for p in range(3):
        c *= g.n[p] + 2 * hw if p < g.dim else 1
    return c


class Context:
    """A HIP context: one device, one stream (default priority)."""

    def __init__(self, device=0):
        self.L = load()
        h = vp()
        check(self.L.scafes_ctx_create(device, C.byref(h)))
        self.h = h
        self._bufs = []

    def close(self):
        if self.h:
            for b in self._bufs:
                if b.ptr:
                    self.L.scafes_free(self.h, b.ptr)
                    b.ptr = None
            self.L.scafes_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(self.L.scafes_ctx_sync(self.h))

    def info(self):
        cu = C.c_int(0)
        name = C.create_string_buffer(64)
        check(self.L.scafes_ctx_info(self.h, C.byref(cu), name, 64))
        return cu.value, name.value.decode()

    def record(self, slot):
        check(self.L.scafes_event_record(self.h, slot))

    def elapsed_ms(self, a, b):
        ms = C.c_float(0)
        check(self.L.scafes_event_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    def alloc(self, count, dtype=np.float64):
        return DeviceArray(self, count, np.dtype(dtype))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        d.write(a)
        return d


class DeviceArray:
    def __init__(self, ctx, count, dtype):
        self.ctx, self.count, self.dtype = ctx, int(count), np.dtype(dtype)
        self.nbytes = self.count * self.dtype.itemsize
        p = vp()
        check(ctx.L.scafes_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p
        ctx._bufs.append(self)

    @property
    def code(self):
        return F64 if self.dtype == np.float64 else F32

    def write(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.count, (a.size, self.count)
        check(self.ctx.L.scafes_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(vp), self.nbytes))

    def read(self, shape=None):
        out = np.empty(self.count, dtype=self.dtype)
        check(self.ctx.L.scafes_memcpy_d2h(self.ctx.h, out.ctypes.data_as(vp), self.ptr, self.nbytes))
        return out.reshape(shape) if shape is not None else out

    def fill(self, value):
        check(self.ctx.L.scafes_fill(self.ctx.h, self.ptr, self.code, self.count, float(value)))

    def free(self):
        if self.ptr:
            check(self.ctx.L.scafes_free(self.ctx.h, self.ptr))
            self.ptr = None


def emfdtd_params(n, d=5, cfln=0.5, src=(10, 10, 10), sink=(10, 10, 14), f_mod=0.0, t0=40.0, spread=8.0,
                  m_a=1.0, m=3.0, k_max=1.0, s_max=0.01222, a_max=1e-3):
    """scafes_emfdtd_params of the EMFDTD example ([PML] / [source] / [sink]
    sections of examples/EMFDTD/testcase_small.ini as defaults)."""
    L = load()
    p = EmfdtdParams()
    p.n, p.d, p.cfln = n, d, cfln
    p.dt = L.scafes_emfdtd_dt(cfln)
    dp = C.POINTER(C.c_double)
    check(L.scafes_emfdtd_tables(d, m_a, m, k_max, s_max, a_max, p.dt, C.cast(p.b, dp), C.cast(p.c, dp),
                                 C.cast(p.ook, dp)))
    for q in range(3):
        p.src[q], p.sink[q] = src[q], sink[q]
    p.f_mod, p.t0, p.spread = f_mod, t0, spread
    return p


def emfdtd_ga_gb(sigma, dt):
    """ga, gb of a host sigma array through the library's long double path."""
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    ga, gb = np.empty_like(sigma), np.empty_like(sigma)
    dp = C.POINTER(C.c_double)
    check(load().scafes_emfdtd_ga_gb(sigma.size, sigma.ctypes.data_as(dp), dt, ga.ctypes.data_as(dp),
                                     gb.ctypes.data_as(dp)))
    return ga, gb
