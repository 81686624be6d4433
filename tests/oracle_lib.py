"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP path and
is never imported by the masterthesis_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

BORDER_KEEP, BORDER_FIELD, BORDER_VALUE = 0, 1, 2
BORDER_NEUMANN, BORDER_CAUCHY, BORDER_RADIATION = 3, 4, 5
KNOWN_NONE, KNOWN_HEAT_TIMELIN, KNOWN_PENNES_TIMELIN = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("hw", C.c_int),
                ("x0", C.c_double * 3), ("x1", C.c_double * 3), ("nlayers", C.c_int)]


class Heat7Coef(C.Structure):
    _fields_ = [("tau", C.c_double), ("coeff_a", C.c_double), ("rhoc_inv", C.c_double),
                ("perf_c", C.c_double), ("perf_w", C.c_double), ("rhob_cb_w", C.c_double),
                ("has_perf", C.c_int), ("has_source", C.c_int), ("lam", C.c_double),
                ("alpha", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double)]


class PennesCoef(C.Structure):
    _fields_ = [("tau", C.c_double), ("cdiff", C.c_double), ("cperf", C.c_double),
                ("csrc", C.c_double), ("t_a", C.c_double)]


def make_grid(n, hw=1, x0=None, x1=None, nlayers=1):
    n = list(n)
    dim = len(n)
    nn = n + [1] * (3 - dim)
    g = Grid()
    g.dim = dim
    for p in range(3):
        g.n[p] = nn[p]
        g.x0[p] = 0.0 if x0 is None or p >= dim else x0[p]
        g.x1[p] = 1.0 if x1 is None or p >= dim else x1[p]
    g.hw = hw
    g.nlayers = nlayers
    return g


def memall_shape(g):
    """numpy shape (z, y, x) of a memAll field (Fortran order = C order reversed)."""
    e = [g.n[p] + 2 * g.hw if p < g.dim else 1 for p in range(3)]
    return (e[2], e[1], e[0])


def nodes_view(g, a):
    """View of the node block (without ghosts) of a memAll array shaped (z,y,x)."""
    h = g.hw
    sl = [slice(h, h + g.n[p]) if p < g.dim else slice(0, 1) for p in range(3)]
    return a[sl[2], sl[1], sl[0]]


def _build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            _build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        gp = C.POINTER(Grid)
        L.oracle_nelems.restype = C.c_long
        L.oracle_nelems.argtypes = [gp]
        L.oracle_gridsize.restype = C.c_double
        L.oracle_gridsize.argtypes = [gp, C.c_int]
        L.oracle_heat7_step.argtypes = [gp, dp, dp, dp, dp, dp, C.c_int, C.c_double, C.POINTER(Heat7Coef)]
        L.oracle_heat7_run.restype = C.c_double
        L.oracle_heat7_run.argtypes = [gp, dp, C.c_int, C.c_double, C.POINTER(Heat7Coef), C.c_int, C.c_int, dp]
        L.oracle_cpml_tables.argtypes = [C.c_int] + [C.c_double] * 6 + [dp, dp, dp]
        L.oracle_emfdtd_dt.restype = C.c_double
        L.oracle_emfdtd_dt.argtypes = [C.c_double]
        L.oracle_emfdtd_ga_gb.argtypes = [C.c_long, dp, C.c_double, dp, dp]
        L.oracle_mod_gauss.restype = C.c_double
        L.oracle_mod_gauss.argtypes = [C.c_double] * 5
        L.oracle_ellipsoid.restype = C.c_double
        L.oracle_ellipsoid.argtypes = [C.c_double] * 10
        L.oracle_emfdtd_step.argtypes = [C.POINTER(EmfdtdParams), C.POINTER(dp), C.c_double, dp]
        L.oracle_emfdtd_run.argtypes = [C.POINTER(EmfdtdParams), C.POINTER(dp), C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_double, dp]
        L.oracle_mri_step.argtypes = [gp, dp, dp, C.POINTER(dp), C.POINTER(MriCoef)]
        L.oracle_tumor_step.argtypes = [gp, dp, dp, C.POINTER(TumorCoef)]
        L.oracle_tumor_fields.argtypes = [gp, C.POINTER(TumorCoef), dp, dp]
        L.oracle_known_timelin.argtypes = [gp, C.c_int, C.c_int, C.c_double, C.POINTER(Heat7Coef), dp, dp, dp]
        L.oracle_heat7var_step.argtypes = [gp, dp, dp, dp, C.c_double, C.c_int, C.c_double]
        L.oracle_heat7var_run.argtypes = [gp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_double, dp]
        L.oracle_heat7var_grad.argtypes = [gp, dp, dp, dp, C.c_double, C.c_int, dp]
        L.oracle_heat7var_grad.restype = C.c_double
        L.oracle_pennes_step.argtypes = [gp, dp, dp, dp, dp, C.POINTER(PennesCoef)]
        L.oracle_pennes_forward_all.argtypes = [gp, dp, dp, dp, C.POINTER(PennesCoef), C.c_int]
        L.oracle_pennes_grad.restype = C.c_double
        L.oracle_pennes_grad.argtypes = [gp, dp, dp, dp, C.POINTER(PennesCoef), C.c_int, dp]
        tb = C.POINTER(TumorBorder)
        L.oracle_pennes_tumor_step.argtypes = [gp, dp, dp, dp, dp, C.POINTER(PennesCoef), tb]
        L.oracle_pennes_tumor_adjoint_step.argtypes = [gp, dp, dp, dp, C.POINTER(PennesCoef), tb]
        L.oracle_pennes_tumor_forward_all.argtypes = [gp, dp, dp, dp, C.POINTER(PennesCoef), tb, C.c_int]
        L.oracle_pennes_tumor_grad.restype = C.c_double
        L.oracle_pennes_tumor_grad.argtypes = [gp, dp, dp, dp, C.POINTER(PennesCoef), tb, C.c_int, dp]
        L.oracle_yee_half_h.argtypes = [C.c_int, C.c_double, dp]
        L.oracle_yee_half_e.argtypes = [C.c_int, C.c_double, dp]
        L.oracle_yee_step.argtypes = [C.c_int, C.c_double, dp]
        L.oracle_yee_energy.restype = C.c_double
        L.oracle_yee_energy.argtypes = [C.c_int, dp, dp]
        L.oracle_yee_eps_half_e.argtypes = [C.c_int, C.c_double, dp, dp]
        L.oracle_yee_eps_step.argtypes = [C.c_int, C.c_double, dp, dp]
        L.oracle_yee_eps_grad.restype = C.c_double
        L.oracle_yee_eps_grad.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_int, dp]
        L.oracle_heat2d_run.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        L.oracle_heat2d_grad.restype = C.c_double
        L.oracle_heat2d_grad.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                         dp, dp, dp]
        L.oracle_linf_nodes.restype = C.c_double
        L.oracle_linf_nodes.argtypes = [gp, dp, dp]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int
        _lib = L
    return _lib


class TumorCoef(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("tau", "k", "rho", "c", "rho_b", "c_pb", "omega_brain", "omega_tumor",
                                          "q_brain", "q_tumor", "t_a", "h", "t_inf", "q_bc", "radius")] + \
               [("center", C.c_double * 3)]


def tumor_coef(tau, center, radius, k=0.6, rho=1040.0, c=3650.0, rho_b=1052.0, c_pb=3800.0, omega_brain=0.004,
               omega_tumor=0.0007, q_brain=25000.0, q_tumor=25000.0, t_a=37.0, h=10.0, t_inf=22.4, q_bc=0.0):
    """Coefficients of the TumorHeatEqnFDM example (Parameters.ini.in defaults)."""
    t = TumorCoef()
    for name, val in dict(tau=tau, k=k, rho=rho, c=c, rho_b=rho_b, c_pb=c_pb, omega_brain=omega_brain,
                          omega_tumor=omega_tumor, q_brain=q_brain, q_tumor=q_tumor, t_a=t_a, h=h, t_inf=t_inf,
                          q_bc=q_bc, radius=radius).items():
        setattr(t, name, val)
    for q in range(3):
        t.center[q] = center[q] if q < len(center) else 0.0
    return t


def tumor_step(g, u, coef):
    un = np.empty_like(u)
    lib().oracle_tumor_step(C.byref(g), ptr(u), ptr(un), C.byref(coef))
    return un


def tumor_fields(g, coef):
    w, q = np.empty(memall_shape(g)), np.empty(memall_shape(g))
    lib().oracle_tumor_fields(C.byref(g), C.byref(coef), ptr(w), ptr(q))
    return w, q


class MriCoef(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("tau", "h", "t_amb", "q_bc", "q_skull", "epsilon", "sigma")]


MRI_FIELDS = ["rho", "c", "lambda", "rho_blood", "c_blood", "omega", "T_blood", "q", "surface"]


def mri_coef(tau, h=10.0, t_amb=22.0, q_bc=0.0, q_skull=0.0, epsilon=0.98, sigma=5.670373e-8):
    m = MriCoef()
    m.tau, m.h, m.t_amb, m.q_bc, m.q_skull, m.epsilon, m.sigma = tau, h, t_amb, q_bc, q_skull, epsilon, sigma
    return m


def mri_step(g, u, mats, coef):
    """mats: 9 stencil-0 node arrays (x fastest) in MRI_FIELDS order."""
    mats = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in mats]
    arr = (C.POINTER(C.c_double) * 9)(*[ptr(x) for x in mats])
    un = np.empty_like(u)
    lib().oracle_mri_step(C.byref(g), ptr(u), ptr(un), arr, C.byref(coef))
    return un


class EmfdtdParams(C.Structure):
    _fields_ = [("n", C.c_int), ("d", C.c_int), ("b", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double)),
                ("ook", C.POINTER(C.c_double)), ("cfln", C.c_double), ("src", C.c_int * 3),
                ("sink", C.c_int * 3)]


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# Convenience wrappers


def heat7_coef(tau, coeff_a=1.0, rhoc_inv=1.0, perf_c=0.0, perf_w=0.0, rhob_cb_w=0.0,
               has_perf=False, has_source=False, lam=1.0, alpha=1.0, epsilon=1.0, sigma=1.0):
    c = Heat7Coef()
    c.tau, c.coeff_a, c.rhoc_inv = tau, coeff_a, rhoc_inv
    c.perf_c, c.perf_w, c.rhob_cb_w = perf_c, perf_w, rhob_cb_w
    c.has_perf, c.has_source = int(has_perf), int(has_source)
    c.lam, c.alpha, c.epsilon, c.sigma = lam, alpha, epsilon, sigma
    return c


def heat7_run(g, u, nsteps, t0, coef, known, border=BORDER_FIELD):
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    linf = np.zeros(nsteps + 1)
    t = lib().oracle_heat7_run(C.byref(g), ptr(u), nsteps, t0, C.byref(coef), known, border, ptr(linf))
    return u, linf, t


def known_timelin(g, known, border, t, coef):
    """F, G, Y (stencil-0, x fastest) of the TimeLinSpaceLin problems at time t."""
    nn = int(np.prod(g.n))
    F, G, Y = np.empty(nn), np.empty(nn), np.empty(nn)
    lib().oracle_known_timelin(C.byref(g), known, border, t, C.byref(coef), ptr(F), ptr(G), ptr(Y))
    return F, G, Y


def heat7_step(g, uo, f_old=None, g_new=None, border=BORDER_VALUE, border_value=0.0, coef=None, g_old=None):
    un = np.empty_like(uo)
    lib().oracle_heat7_step(C.byref(g), ptr(uo), ptr(un), ptr(f_old), ptr(g_old), ptr(g_new), border,
                            border_value, C.byref(coef))
    return un


def heat7var_step(g, uo, kappa, tau, border=BORDER_VALUE, border_value=0.0):
    un = np.empty_like(uo)
    lib().oracle_heat7var_step(C.byref(g), ptr(uo), ptr(un), ptr(kappa), tau, border, border_value)
    return un


def heat7var_run(g, u, kappa, tau, nsteps, border=BORDER_VALUE, border_value=0.0):
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    scratch = np.empty_like(u)
    lib().oracle_heat7var_run(C.byref(g), ptr(u), ptr(kappa), tau, nsteps, border,
                              border_value, ptr(scratch))
    return u


def heat7var_forward_all(g, u0, kappa, tau, nsteps, border=BORDER_VALUE, border_value=0.0):
    """u^0..u^nsteps of the variable-kappa step, shape (nsteps+1,) + u0.shape."""
    st = np.empty((nsteps + 1,) + u0.shape)
    st[0] = u0
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    for n in range(nsteps):
        st[n + 1] = heat7var_step(g, np.ascontiguousarray(st[n]), kap, tau, border, border_value)
    return st


def heat7var_grad(g, states, kappa, target, tau, nsteps):
    """(J, dJ/dkappa) for J = 0.5 sum_inner (u^N - target)^2."""
    grad = np.zeros(states[0].shape)
    J = lib().oracle_heat7var_grad(C.byref(g), ptr(np.ascontiguousarray(states)),
                                   ptr(np.ascontiguousarray(kappa, dtype=np.float64)),
                                   ptr(np.ascontiguousarray(target, dtype=np.float64)), tau, nsteps, ptr(grad))
    return J, grad


def pennes_coef(tau, cdiff, cperf, csrc, t_a):
    c = PennesCoef()
    c.tau, c.cdiff, c.cperf, c.csrc, c.t_a = tau, cdiff, cperf, csrc, t_a
    return c


def pennes_forward_all(g, t0, w, q, coef, nsteps):
    ne = t0.size
    st = np.empty((nsteps + 1,) + t0.shape)
    st[0] = t0
    lib().oracle_pennes_forward_all(C.byref(g), ptr(st), ptr(w), ptr(q), C.byref(coef), nsteps)
    assert st[0].size == ne
    return st


def pennes_grad(g, states, w, target, coef, nsteps):
    grad = np.zeros_like(w)
    J = lib().oracle_pennes_grad(C.byref(g), ptr(states), ptr(w), ptr(target), C.byref(coef),
                                 nsteps, ptr(grad))
    return J, grad


class TumorBorder(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k", "h", "t_inf", "q_bc")]


def tumor_border(k, h, t_inf, q_bc):
    b = TumorBorder()
    b.k, b.h, b.t_inf, b.q_bc = k, h, t_inf, q_bc
    return b


def pennes_tumor_step(g, u, w, q, coef, border):
    un = np.empty_like(u)
    lib().oracle_pennes_tumor_step(C.byref(g), ptr(u), ptr(un), ptr(w), ptr(q), C.byref(coef), C.byref(border))
    return un


def pennes_tumor_adjoint_step(g, lam, w, coef, border):
    out = np.zeros_like(lam)
    lib().oracle_pennes_tumor_adjoint_step(C.byref(g), ptr(lam), ptr(out), ptr(w), C.byref(coef), C.byref(border))
    return out


def pennes_tumor_forward_all(g, t0, w, q, coef, border, nsteps):
    st = np.empty((nsteps + 1,) + t0.shape)
    st[0] = t0
    lib().oracle_pennes_tumor_forward_all(C.byref(g), ptr(st), ptr(w), ptr(q), C.byref(coef), C.byref(border),
                                          nsteps)
    return st


def pennes_tumor_grad(g, states, w, target, coef, border, nsteps):
    grad = np.zeros_like(w)
    J = lib().oracle_pennes_tumor_grad(C.byref(g), ptr(states), ptr(w), ptr(target), C.byref(coef),
                                       C.byref(border), nsteps, ptr(grad))
    return J, grad


def yee_alloc(ncell):
    m = ncell + 1
    return np.zeros((6, m, m, m))


def yee_step(ncell, s, f, nsteps=1):
    for _ in range(nsteps):
        lib().oracle_yee_step(ncell, s, ptr(f))
    return f


def yee_half_h(ncell, s, f):
    lib().oracle_yee_half_h(ncell, s, ptr(f))


def yee_half_e(ncell, s, f):
    lib().oracle_yee_half_e(ncell, s, ptr(f))


def yee_energy(ncell, f, h_prev):
    return lib().oracle_yee_energy(ncell, ptr(f), ptr(np.ascontiguousarray(h_prev)))


def yee_eps_step(ncell, s, eps, f, nsteps=1):
    """f: (6, N+1, N+1, N+1) state [field, k, j, i]; eps: (N+1)^3."""
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    for _ in range(nsteps):
        lib().oracle_yee_eps_step(ncell, s, ptr(eps), ptr(f))
    return f


def yee_eps_forward_all(ncell, s, eps, f0, nsteps):
    """(nsteps+1, 6, N+1, N+1, N+1): the states 0..nsteps."""
    st = np.empty((nsteps + 1,) + f0.shape)
    st[0] = f0
    for n in range(nsteps):
        st[n + 1] = st[n]
        yee_eps_step(ncell, s, eps, st[n + 1])
    return st


def yee_eps_grad(ncell, s, eps, states, target, nsteps):
    """(J, dJ/deps) with target the three E arrays (3, N+1, N+1, N+1)."""
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    grad = np.zeros(eps.shape)
    J = lib().oracle_yee_eps_grad(ncell, s, ptr(eps), ptr(np.ascontiguousarray(states)),
                                  ptr(np.ascontiguousarray(target, dtype=np.float64)), nsteps, ptr(grad))
    return J, grad


def yee_emask(ncell):
    """(3, N+1, N+1, N+1) bool: the E entries the E half step updates
    (tangential E on the PEC walls and the padding entries stay fixed)."""
    m = ncell + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    inner = lambda a: (a >= 1) & (a < ncell)
    return np.stack([(i < ncell) & inner(j) & inner(k), (j < ncell) & inner(i) & inner(k),
                     (k < ncell) & inner(i) & inner(j)])


def heat2d_run(n, kappa, tau, nsteps, u0):
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    up = np.empty_like(u)
    lib().oracle_heat2d_run(n, 1.0 / (n - 1), kappa, tau, nsteps, ptr(u), ptr(up))
    return u, up


def heat2d_grad(n, kappa, tau, nsteps, u0, target):
    d = C.c_double(0.0)
    J = lib().oracle_heat2d_grad(n, 1.0 / (n - 1), kappa, tau, nsteps,
                                 ptr(np.ascontiguousarray(u0, dtype=np.float64)),
                                 ptr(np.ascontiguousarray(target, dtype=np.float64)), C.byref(d))
    return J, d.value


# ---------------------------------------------------------------------------
# EMFDTD example (Yee leapfrog + CPML + conductive media)

EM_FIELDS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ix", "Iy", "Iz",
             "Psi_ExDy", "Psi_ExDz", "Psi_EyDx", "Psi_EyDz", "Psi_EzDx", "Psi_EzDy",
             "Psi_HxDy", "Psi_HxDz", "Psi_HyDx", "Psi_HyDz", "Psi_HzDx", "Psi_HzDy", "ga", "gb"]
EM = {name: i for i, name in enumerate(EM_FIELDS)}
EM_STATE = 24  # fields updated by the step (ga, gb are read only)


def cpml_tables(d, m_a=1.0, m=3.0, k_max=1.0, s_max=0.01222, a_max=1e-3, dt=None):
    """CPMLParams tables (b, c, 1/kappa), d+1 entries, index 0 = outermost."""
    if dt is None:
        dt = emfdtd_dt(0.5)
    b, c, ook = np.empty(d + 1), np.empty(d + 1), np.empty(d + 1)
    lib().oracle_cpml_tables(d, m_a, m, k_max, s_max, a_max, dt, ptr(b), ptr(c), ptr(ook))
    return b, c, ook


def emfdtd_dt(cfln=0.5):
    return lib().oracle_emfdtd_dt(cfln)


def ga_gb(sigma, dt):
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    ga, gb = np.empty_like(sigma), np.empty_like(sigma)
    lib().oracle_emfdtd_ga_gb(sigma.size, ptr(sigma), dt, ptr(ga), ptr(gb))
    return ga, gb


def mod_gauss(t, dt, f_mod=0.0, t0=40.0, spread=8.0):
    return lib().oracle_mod_gauss(t, dt, f_mod, t0, spread)


def ellipsoid_sigma(n, center, radii, inside=3.546e7):
    """SIGMA of EMFDTDng::initInner (EMFDTD.hpp:258-276): the indicator at the
    node indices, (n,n,n) array indexed [k, j, i]."""
    out = np.empty((n, n, n))
    L = lib()
    for k in range(n):
        for j in range(n):
            for i in range(n):
                out[k, j, i] = L.oracle_ellipsoid(float(i), float(j), float(k), *map(float, center),
                                                  *map(float, radii), inside)
    return out


class Emfdtd:
    """State of the EMFDTD example on an n^3 grid: f[26, n, n, n] (x fastest)."""

    def __init__(self, n, d=5, cfln=0.5, src=(10, 10, 10), sink=(10, 10, 14), sigma=None, tables=None, **pml):
        self.n, self.d, self.cfln = n, d, cfln
        self.dt = emfdtd_dt(cfln)
        b, c, ook = tables if tables is not None else cpml_tables(d, dt=self.dt, **pml)
        self.b, self.c, self.ook = (np.ascontiguousarray(x, dtype=np.float64) for x in (b, c, ook))
        self.f = np.zeros((len(EM_FIELDS), n, n, n))
        sig = np.zeros((n, n, n)) if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        ga, gb = ga_gb(sig.ravel(), self.dt)
        self.f[EM["ga"]] = ga.reshape(n, n, n)
        self.f[EM["gb"]] = gb.reshape(n, n, n)
        self.p = EmfdtdParams()
        self.p.n, self.p.d, self.p.cfln = n, d, cfln
        self.p.b, self.p.c, self.p.ook = ptr(self.b), ptr(self.c), ptr(self.ook)
        for q in range(3):
            self.p.src[q], self.p.sink[q] = src[q], sink[q]

    def _fp(self):
        arr = (C.POINTER(C.c_double) * len(EM_FIELDS))()
        for i in range(len(EM_FIELDS)):
            arr[i] = self.f[i].ctypes.data_as(C.POINTER(C.c_double))
        return arr

    def step(self, src_value):
        out = C.c_double(0.0)
        lib().oracle_emfdtd_step(C.byref(self.p), self._fp(), src_value, C.byref(out))
        return out.value

    def run(self, nsteps, t_start=0.0, f_mod=0.0, t0=40.0, spread=8.0):
        sink = np.zeros(max(nsteps, 1))
        lib().oracle_emfdtd_run(C.byref(self.p), self._fp(), nsteps, t_start, f_mod, t0, spread, ptr(sink))
        return sink[:nsteps]

    def __getitem__(self, name):
        return self.f[EM[name]]
