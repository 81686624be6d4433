"""Pins of the CPU oracle against what the reference's tests and the
mathematics fix.  Each test is chosen so a plausible slip in the oracle
(a dropped term, a wrong sign or index, F_new instead of F_old, G_old
instead of G_new, a wrong face average, a wrong curl sign) fails it.

 * TimeLinSpaceLin heat / Pennes (tests/system/{HeatEqnFDM,PennesBioheatEqnFDM}/
   {Dirichlet,Neumann,Cauchy,Radiation}/TimeLinSpaceLin): y = (1+t)(1+sum x)
   is reproduced exactly by explicit Euler + 7-point stencil and the
   one-sided flux borders, so every "Local Linf" must be round-off.
 * Discrete sine eigenmode: u^n = (1 - tau*lambda_h)^n prod sin(pi x).
 * Var-kappa: symmetry of the discrete operator, reduction to constant
   kappa, face-by-face assembled operator.
 * Pennes field form: closed-form steady state, reduction to heat, adjoint
   gradient against central finite differences.
 * Yee: PEC cavity eigenmodes with the exact Yee dispersion relation,
   discrete energy invariant, discrete Gauss law.
 * 2-D heat: eigenmode closed form, adjoint gradient vs finite differences.
"""
import math

import numpy as np
import pytest

import oracle_lib as O


# ---------------------------------------------------------------------------
# TimeLinSpaceLin system problems (exact discrete solutions)

@pytest.mark.parametrize("dims", [(9,), (8, 8), (8, 8, 8), (12, 10, 9)])
def test_heat_timelin_exact(dims):
    g = O.make_grid(dims)
    h2 = sum(1.0 / O.lib().oracle_gridsize(g, p) ** 2 for p in range(len(dims)))
    nsteps = 60
    tau = 0.9 / (2.0 * h2)  # stability condition of main3D.cpp:45-53
    coef = O.heat7_coef(tau, coeff_a=1.0, rhoc_inv=1.0, has_source=True)
    u = np.zeros(O.memall_shape(g))
    u, linf, t = O.heat7_run(g, u, nsteps, 0.0, coef, O.KNOWN_HEAT_TIMELIN)
    assert linf.max() < 1e-13 * (1 + nsteps * tau) * 10
    assert abs(t - nsteps * tau) < 1e-14 * nsteps
    # the final field equals the exact solution at t_N
    x = [np.linspace(0, 1, n) for n in dims]
    mesh = np.meshgrid(*x, indexing="ij")
    y = (1.0 + sum(mesh)) * (1.0 + nsteps * tau)
    got = O.nodes_view(g, u).reshape(tuple(reversed(dims))).T
    np.testing.assert_allclose(got, y, rtol=1e-13, atol=0)


@pytest.mark.parametrize("form", [O.KNOWN_HEAT_TIMELIN, O.KNOWN_PENNES_TIMELIN])
@pytest.mark.parametrize("border", [O.BORDER_FIELD, O.BORDER_NEUMANN, O.BORDER_CAUCHY, O.BORDER_RADIATION])
def test_known_timelin_fields_closed_form(form, border):
    """The known-field evaluator on its own, on a shifted 3-D domain with two
    border layers: Y = (1+t)(1+sum x) everywhere, F = (1+sum x) [+ w (1+t)(1+sum x)]
    (zero on Dirichlet border nodes), G = Y on Dirichlet border nodes and on
    flux-border nodes with an index 0, the flux datum on the far faces (the
    value the one-sided border update needs for y to stay exact: for
    Neumann -lambda dy/dn = -lambda (1+t))."""
    dims, x0, x1, nl, t = (7, 6, 5), (0.5, -1.0, 0.0), (2.0, 1.5, 0.8), 2, 0.75
    g = O.make_grid(dims, hw=0, x0=x0, x1=x1, nlayers=nl)
    lam, alpha, eps, sig, w = 1.3, 0.6, 0.9, 0.4, 0.7
    coef = O.heat7_coef(1e-3, rhob_cb_w=w, lam=lam, alpha=alpha, epsilon=eps, sigma=sig)
    F, G, Y = (a.reshape(tuple(reversed(dims))).T for a in O.known_timelin(g, form, border, t, coef))
    ax = [np.linspace(x0[p], x1[p], dims[p]) for p in range(3)]
    X = np.meshgrid(*ax, indexing="ij")
    s = 1.0 + X[0] + X[1] + X[2]
    y = s * (1.0 + t)
    idx = np.meshgrid(*(np.arange(n) for n in dims), indexing="ij")
    bord = np.zeros(dims, bool)
    for p in range(3):
        bord |= (idx[p] < nl) | (idx[p] > dims[p] - 1 - nl)
    near = (idx[0] == 0) | (idx[1] == 0) | (idx[2] == 0)
    f = s + (w * y if form == O.KNOWN_PENNES_TIMELIN else 0.0)
    np.testing.assert_allclose(Y, y, rtol=1e-14)
    if border == O.BORDER_FIELD:
        np.testing.assert_allclose(F, np.where(bord, 0.0, f), rtol=1e-14)
        np.testing.assert_allclose(G, np.where(bord, y, 0.0), rtol=1e-14)
        return
    np.testing.assert_allclose(F, f, rtol=1e-14)
    datum = {O.BORDER_NEUMANN: -lam * (1.0 + t) + 0.0 * s,
             O.BORDER_CAUCHY: alpha * y + lam * (1.0 + t),
             O.BORDER_RADIATION: sig * eps * y ** 4 + lam * (1.0 + t)}[border]
    want = np.where(bord, np.where(near, y, datum), 0.0)
    np.testing.assert_allclose(G, want, rtol=1e-14)


def test_heat_timelin_detects_missing_source():
    """A run without the source term drifts away from the exact solution."""
    g = O.make_grid((8, 8, 8))
    coef = O.heat7_coef(1e-3, has_source=False)
    _, linf, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), 20, 0.0, coef, O.KNOWN_HEAT_TIMELIN)
    assert linf.max() > 1e-3


@pytest.mark.parametrize("dims", [(10,), (8, 9), (8, 8, 8)])
def test_pennes_timelin_exact(dims):
    g = O.make_grid(dims)
    h2 = sum(1.0 / O.lib().oracle_gridsize(g, p) ** 2 for p in range(len(dims)))
    tau = 0.8 / (2.0 * h2 + 1.0)
    coef = O.heat7_coef(tau, coeff_a=1.0, rhoc_inv=1.0, perf_c=1.0, perf_w=1.0, rhob_cb_w=1.0,
                        has_perf=True, has_source=True)
    _, linf, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), 40, 0.0, coef, O.KNOWN_PENNES_TIMELIN)
    assert linf.max() < 1e-12


def test_pennes_timelin_needs_f_old():
    """Pennes F depends on t: evaluating the source at t_n instead of t_{n-1}
    (a plausible slip) breaks exactness; the exact-solution check sees it.
    Emulated by shifting t0 by one step for the known fields only."""
    g = O.make_grid((8, 8, 8))
    tau = 1e-3
    coef = O.heat7_coef(tau, perf_c=1.0, perf_w=1.0, rhob_cb_w=1.0, has_perf=True, has_source=True)
    _, linf_ok, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), 30, 0.0, coef, O.KNOWN_PENNES_TIMELIN)
    assert linf_ok.max() < 1e-13
    # perturbing W in the update only (not in F) must show a large error
    bad = O.heat7_coef(tau, perf_c=1.0, perf_w=1.1, rhob_cb_w=1.0, has_perf=True, has_source=True)
    _, linf_bad, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), 30, 0.0, bad, O.KNOWN_PENNES_TIMELIN)
    assert linf_bad.max() > 1e-4


# ---------------------------------------------------------------------------
# Flux border classes of the reference's system tests (Neumann, Cauchy,
# radiation; tests/system/*/{Neumann,Cauchy,Radiation}/TimeLinSpaceLin): the
# one-sided border update with the flux datum G_old keeps y = (1+t)(1+sum x)
# exact, so every "Local Linf" must be round-off.

FLUX_CASES = [("heat", O.BORDER_NEUMANN), ("heat", O.BORDER_CAUCHY), ("pennes", O.BORDER_NEUMANN),
              ("pennes", O.BORDER_CAUCHY), ("pennes", O.BORDER_RADIATION)]


def flux_coef(form, tau, **kw):
    if form == "pennes":
        return O.heat7_coef(tau, perf_c=1.0, perf_w=1.0, rhob_cb_w=1.0, has_perf=True, has_source=True, **kw)
    return O.heat7_coef(tau, has_source=True, **kw)


@pytest.mark.parametrize("form,border", FLUX_CASES)
@pytest.mark.parametrize("dims", [(11,), (9, 8), (8, 9, 7)])
def test_flux_border_timelin_exact(dims, form, border):
    g = O.make_grid(dims)
    h2 = sum(1.0 / O.lib().oracle_gridsize(g, p) ** 2 for p in range(len(dims)))
    nsteps = 40
    tau = 0.4 / (2.0 * h2 + 1.0)
    if border == O.BORDER_RADIATION:
        # the linearised radiation term adds 8 eps sigma u^3 / h to the
        # stiffness (u <= 1 + dim here)
        hmin = min(O.lib().oracle_gridsize(g, p) for p in range(len(dims)))
        tau = 0.4 / (2.0 * h2 + 1.0 + 8.0 * (1.0 + len(dims)) ** 3 / hmin)
    known = O.KNOWN_PENNES_TIMELIN if form == "pennes" else O.KNOWN_HEAT_TIMELIN
    u, linf, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), nsteps, 0.0, flux_coef(form, tau), known, border)
    assert linf.max() < 1e-12
    # the far faces really are updated by the flux rule (not kept)
    x = [np.linspace(0, 1, n) for n in dims]
    y = (1.0 + sum(np.meshgrid(*x, indexing="ij"))) * (1.0 + nsteps * tau)
    got = O.nodes_view(g, u).reshape(tuple(reversed(dims))).T
    np.testing.assert_allclose(got[(-1,) * len(dims)], y[(-1,) * len(dims)], rtol=1e-12)


def _timelin_known(g, dims, t, form, border, lam=1.0, alpha=1.0, eps_sigma=1.0):
    """F, G (stencil-0, (z,y,x)) of the flux-border TimeLinSpaceLin tests and
    the exact state y, computed independently of the oracle."""
    x = [np.linspace(0, 1, n) for n in dims]
    mesh = np.meshgrid(*reversed(x), indexing="ij")  # (z, y, x) order, axis 0 = last dim
    s = 1.0 + sum(mesh)
    y = s * (1.0 + t)
    F = s.copy()
    if form == "pennes":
        F = F + y
    idx = np.meshgrid(*(np.arange(n) for n in reversed(dims)), indexing="ij")
    near = np.zeros(s.shape, bool)
    for a in idx:
        near |= (a == 0)
    if border == O.BORDER_NEUMANN:
        flux = np.full(s.shape, -lam * (1.0 + t))
    elif border == O.BORDER_CAUCHY:
        flux = alpha * y + lam * (1.0 + t)
    else:
        flux = eps_sigma * y ** 4 + lam * (1.0 + t)
    G = np.where(near, y, flux)
    return F, G, y


@pytest.mark.parametrize("form,border", FLUX_CASES)
def test_flux_border_one_step_sensitivity(form, border):
    """One oracle step from the exact state with independently evaluated
    known fields lands on y(t+tau); a flux datum with the wrong sign, or G_new
    used where the rule reads G_old, does not."""
    dims = (7, 8, 6)
    g = O.make_grid(dims)
    tau, t = 2e-3, 0.3
    F, Go, y0 = _timelin_known(g, dims, t, form, border)
    _, Gn, y1 = _timelin_known(g, dims, t + tau, form, border)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = y0.reshape(O.nodes_view(g, u).shape)
    c = flux_coef(form, tau)
    un = O.heat7_step(g, u, f_old=F.ravel(), g_new=Gn.ravel(), g_old=Go.ravel(), border=border, coef=c)
    err = np.abs(O.nodes_view(g, un).reshape(y1.shape) - y1).max()
    assert err < 1e-12
    for bad in (-Go, Gn):
        un = O.heat7_step(g, u, f_old=F.ravel(), g_new=Gn.ravel(), g_old=bad.ravel(), border=border, coef=c)
        assert np.abs(O.nodes_view(g, un).reshape(y1.shape) - y1).max() > 1e-6


# ---------------------------------------------------------------------------
# Sine eigenmode (config: heat 64^3 plain method)

def sine_init(g):
    n = [g.n[p] for p in range(g.dim)]
    x = [np.array([O.lib().oracle_gridsize(g, p) * i for i in range(n[p])]) for p in range(g.dim)]
    mesh = np.meshgrid(*[np.sin(np.pi * xi) for xi in x], indexing="ij")
    prod = np.ones_like(mesh[0])
    for m in mesh:
        prod = prod * m
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = prod.T.reshape(O.nodes_view(g, u).shape)
    return u, prod


@pytest.mark.parametrize("n", [16, 24])
def test_sine_mode_closed_form(n):
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 6.0
    nsteps = 50
    u0, prod = sine_init(g)
    coef = O.heat7_coef(tau, coeff_a=1.0)
    u, _, _ = O.heat7_run(g, u0, nsteps, 0.0, coef, O.KNOWN_NONE)
    lam = 3 * 4.0 / h ** 2 * math.sin(math.pi * h / 2) ** 2
    factor = (1.0 - tau * lam) ** nsteps
    got = O.nodes_view(g, u).T
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(got[inner], factor * prod[inner], rtol=0, atol=5e-14)
    # border nodes are Dirichlet zero after the first step
    assert np.all(got[0] == 0.0) and np.all(got[-1] == 0.0)


# ---------------------------------------------------------------------------
# Variable-coefficient heat

def kappa_field(g, seed=1):
    rng = np.random.default_rng(seed)
    return 1.0 + 0.5 * rng.random(O.memall_shape(g))


def test_heat7var_reduces_to_constant():
    g = O.make_grid((11, 10, 9))
    rng = np.random.default_rng(0)
    u = rng.random(O.memall_shape(g))
    kap = np.full(O.memall_shape(g), 0.7)
    tau = 1e-3
    a = O.heat7var_step(g, u, kap, tau, O.BORDER_KEEP)
    b = O.heat7_step(g, u, coef=O.heat7_coef(tau, coeff_a=0.7), border=O.BORDER_KEEP)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)


def test_heat7var_operator_symmetric():
    """The conservative face-average form is a symmetric operator on the
    inner nodes (zero Dirichlet border); a non-conservative kappa_C*lap(u)
    or a one-sided face coefficient is not."""
    g = O.make_grid((9, 8, 7))
    kap = kappa_field(g)
    tau = 1e-3
    rng = np.random.default_rng(3)
    inner = np.zeros(O.memall_shape(g), bool)
    O.nodes_view(g, inner)[1:-1, 1:-1, 1:-1] = True

    def A(v):
        v = np.where(inner, v, 0.0)
        return np.where(inner, O.heat7var_step(g, v, kap, tau, O.BORDER_VALUE, 0.0), 0.0)

    v = rng.standard_normal(O.memall_shape(g))
    w = rng.standard_normal(O.memall_shape(g))
    lhs = np.sum(A(v) * np.where(inner, w, 0))
    rhs = np.sum(np.where(inner, v, 0) * A(w))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # and it is not the identity
    assert np.abs(A(v) - np.where(inner, v, 0)).max() > 1e-6


def test_heat7var_face_assembly():
    """Assemble the update face by face (each face flux added to one node and
    subtracted from the other) and compare with the node-centred oracle."""
    g = O.make_grid((7, 6, 5))
    kap = kappa_field(g, 5)
    rng = np.random.default_rng(4)
    u = rng.random(O.memall_shape(g))
    tau = 2e-4
    got = O.heat7var_step(g, u, kap, tau, O.BORDER_KEEP)
    acc = np.zeros_like(u)
    hs = [O.lib().oracle_gridsize(g, p) for p in range(3)]
    U = O.nodes_view(g, u)  # (z,y,x) incl. all nodes
    K = O.nodes_view(g, kap)
    for ax, p in ((2, 0), (1, 1), (0, 2)):  # numpy axis, grid axis
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(0, -1)
        sl_hi[ax] = slice(1, None)
        flux = 0.5 * (K[tuple(sl_lo)] + K[tuple(sl_hi)]) * (U[tuple(sl_hi)] - U[tuple(sl_lo)]) / hs[p] ** 2
        d = np.zeros_like(U)
        d[tuple(sl_lo)] += flux
        d[tuple(sl_hi)] -= flux
        acc_nodes = O.nodes_view(g, acc)
        acc_nodes += d
    expect = u + tau * acc
    inner = (slice(2, -2),) * 3  # memAll indices of inner nodes
    np.testing.assert_allclose(got[inner], expect[inner], rtol=1e-12, atol=1e-14)


def test_heat7var_sine_decay_constant_kappa():
    n = 16
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    kap = np.full(O.memall_shape(g), 1.5)
    tau = 0.9 * h * h / (6 * 1.5)
    u0, prod = sine_init(g)
    u = O.heat7var_run(g, u0, kap, tau, 40)
    lam = 3 * 4.0 / h ** 2 * math.sin(math.pi * h / 2) ** 2
    f = (1 - tau * 1.5 * lam) ** 40
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(O.nodes_view(g, u).T[inner], f * prod[inner], atol=5e-14)


# ---------------------------------------------------------------------------
# Pennes with perfusion / source fields

def test_pennes_steady_state():
    g = O.make_grid((8, 9, 10))
    w, q, ta = 0.8, 3.0, 37.0
    c = O.pennes_coef(tau=1e-3, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=ta)
    T = np.full(O.memall_shape(g), ta + q / w)
    W = np.full_like(T, w)
    Q = np.full_like(T, q)
    st = O.pennes_forward_all(g, T, W, Q, c, 20)
    np.testing.assert_allclose(st[-1], T, rtol=1e-14)
    # a sign slip in either the perfusion or the source term moves the state
    c2 = O.pennes_coef(tau=1e-3, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=ta + 1.0)
    st2 = O.pennes_forward_all(g, T, W, Q, c2, 5)
    assert np.abs(st2[-1] - T).max() > 1e-4


def test_pennes_reduces_to_heat():
    g = O.make_grid((9, 8, 7))
    rng = np.random.default_rng(7)
    T = rng.random(O.memall_shape(g))
    c = O.pennes_coef(tau=2e-4, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=37.0)
    st = O.pennes_forward_all(g, T, np.zeros_like(T), np.zeros_like(T), c, 1)
    b = O.heat7_step(g, T, coef=O.heat7_coef(2e-4, coeff_a=0.5), border=O.BORDER_KEEP)
    np.testing.assert_array_equal(st[1], b)


def _pennes_setup(n=8, nsteps=25, seed=11):
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(seed)
    w = 0.5 * np.exp(0.3 * rng.standard_normal(O.memall_shape(g)))
    x = np.linspace(0, 1, n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    q = np.zeros(O.memall_shape(g))
    O.nodes_view(g, q)[...] = 50.0 * np.exp(-((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) / (2 * .15 ** 2))
    tau = 0.9 / (6 * 0.5 / h ** 2 + w.max())
    c = O.pennes_coef(tau=tau, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=37.0)
    T0 = np.full(O.memall_shape(g), 37.0)
    w_pert = w * (1 + 0.2 * rng.standard_normal(w.shape))
    target = O.pennes_forward_all(g, T0, w_pert, q, c, nsteps)[-1]
    return g, w, q, c, T0, target, nsteps


def test_pennes_adjoint_vs_finite_differences():
    g, w, q, c, T0, target, N = _pennes_setup()

    def J(wv):
        TN = O.pennes_forward_all(g, T0, wv, q, c, N)[-1]
        inner = np.zeros(TN.shape, bool)
        O.nodes_view(g, inner)[1:-1, 1:-1, 1:-1] = True
        return 0.5 * np.sum(((TN - target) * inner) ** 2)

    st = O.pennes_forward_all(g, T0, w, q, c, N)
    J0, grad = O.pennes_grad(g, st, w, target, c, N)
    assert abs(J0 - J(w)) < 1e-12 * abs(J0)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(w.shape)
    eps = 1e-4
    fd = (J(w + eps * v) - J(w - eps * v)) / (2 * eps)
    ad = np.sum(grad * v)
    assert abs(fd - ad) < 1e-6 * abs(ad)
    # pointwise at a few inner nodes
    for (k, j, i) in [(3, 3, 3), (4, 2, 5), (1, 1, 1)]:
        e = np.zeros_like(w)
        O.nodes_view(g, e)[k, j, i] = 1.0
        fd = (J(w + 1e-3 * e) - J(w - 1e-3 * e)) / 2e-3
        assert abs(fd - O.nodes_view(g, grad)[k, j, i]) < 1e-6 * abs(fd) + 1e-12


# ---------------------------------------------------------------------------
# Yee FDTD (PEC cavity)

def _yee_mode(N, axis, p, q):
    """TM mode uniform along `axis`: E_axis = sin(pi p a/N) sin(pi q b/N)."""
    f = O.yee_alloc(N)
    idx = np.arange(N + 1)
    sa = np.sin(np.pi * p * idx / N)
    sb = np.sin(np.pi * q * idx / N)
    # array layout f[c, k, j, i]
    if axis == 2:    # Ez(i,j): depends on i (x) and j (y); valid k < N
        f[2, :N, :, :] = np.outer(sb, sa)[None, :, :]
    elif axis == 0:  # Ex(j,k): valid i < N
        f[0, :, :, :N] = np.outer(sb, sa)[:, :, None]
    else:            # Ey(i,k): valid j < N
        f[1, :, :N, :] = np.outer(sb, sa)[:, None, :]
    return f


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_yee_cavity_mode_dispersion(axis):
    N, s, p, q, nsteps = 12, 0.5, 1, 2, 40
    f = _yee_mode(N, axis, p, q)
    mode = f[axis].copy()
    O.yee_step(N, s, f, nsteps)
    th = 2 * math.asin(s * math.sqrt(math.sin(math.pi * p / (2 * N)) ** 2 +
                                     math.sin(math.pi * q / (2 * N)) ** 2))
    a = math.cos(nsteps * th) - math.tan(th / 2) * math.sin(nsteps * th)
    np.testing.assert_allclose(f[axis], a * mode, atol=1e-13)
    # the other E components stay zero
    for c in range(3):
        if c != axis:
            assert np.abs(f[c]).max() == 0.0


def test_yee_energy_and_gauss_law():
    N, s = 10, 0.5
    rng = np.random.default_rng(5)
    f = O.yee_alloc(N)
    # random E with zero tangential components on the PEC walls
    f[0, 1:N, 1:N, :N] = rng.standard_normal((N - 1, N - 1, N))
    f[1, 1:N, :N, 1:N] = rng.standard_normal((N - 1, N, N - 1))
    f[2, :N, 1:N, 1:N] = rng.standard_normal((N, N - 1, N - 1))

    def div_e(f):
        Ex, Ey, Ez = f[0], f[1], f[2]
        d = np.zeros((N + 1,) * 3)
        d[1:N, 1:N, 1:N] = ((Ex[1:N, 1:N, 1:N] - Ex[1:N, 1:N, 0:N - 1]) +
                            (Ey[1:N, 1:N, 1:N] - Ey[1:N, 0:N - 1, 1:N]) +
                            (Ez[1:N, 1:N, 1:N] - Ez[0:N - 1, 1:N, 1:N]))
        return d

    d0 = div_e(f)
    energies = []
    for _ in range(30):
        hprev = f[3:].copy()
        O.yee_half_h(N, s, f)
        energies.append(O.yee_energy(N, f, hprev))
        O.yee_half_e(N, s, f)
    energies = np.array(energies)
    assert np.abs(energies - energies[0]).max() < 1e-12 * energies[0]
    np.testing.assert_allclose(div_e(f), d0, atol=1e-12)


# ---------------------------------------------------------------------------
# 2-D heat ensemble member

def test_heat2d_sine_closed_form():
    n, kappa, nsteps = 33, 0.6, 80
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 4.0
    x = np.linspace(0, 1, n)
    u0 = np.outer(np.sin(np.pi * x), np.sin(np.pi * x))
    u0[0, :] = u0[-1, :] = u0[:, 0] = u0[:, -1] = 0.0
    u, _ = O.heat2d_run(n, kappa, tau, nsteps, u0)
    lam = 2 * 4 / h ** 2 * math.sin(math.pi * h / 2) ** 2
    np.testing.assert_allclose(u, (1 - tau * kappa * lam) ** nsteps * u0, atol=1e-14)


def test_heat2d_gradient_fd_and_identity():
    n, kappa, nsteps = 20, 0.4, 60
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 4.0
    rng = np.random.default_rng(9)
    u0 = rng.random((n, n))
    u0[0, :] = u0[-1, :] = u0[:, 0] = u0[:, -1] = 0.0
    target = O.heat2d_run(n, kappa * 1.3, tau, nsteps, u0)[0]
    J, g = O.heat2d_grad(n, kappa, tau, nsteps, u0, target)

    def Jf(k):
        u = O.heat2d_run(n, k, tau, nsteps, u0)[0]
        return 0.5 * np.sum((u[1:-1, 1:-1] - target[1:-1, 1:-1]) ** 2)

    eps = 1e-5
    fd = (Jf(kappa + eps) - Jf(kappa - eps)) / (2 * eps)
    assert abs(J - Jf(kappa)) < 1e-13 * J
    assert abs(fd - g) < 1e-7 * abs(g)
    # commuting-operator identity used by the batched kernel:
    # dJ/dkappa = N * tau * <u^N - target, L u^{N-1}>
    uN, uNm1 = O.heat2d_run(n, kappa, tau, nsteps, u0)
    L = np.zeros_like(uNm1)
    L[1:-1, 1:-1] = ((uNm1[1:-1, :-2] + uNm1[1:-1, 2:] - 2 * uNm1[1:-1, 1:-1]) +
                     (uNm1[:-2, 1:-1] + uNm1[2:, 1:-1] - 2 * uNm1[1:-1, 1:-1])) / h ** 2
    ident = nsteps * tau * np.sum((uN - target)[1:-1, 1:-1] * L[1:-1, 1:-1])
    assert abs(ident - g) < 1e-10 * abs(g)
