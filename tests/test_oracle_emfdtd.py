"""Pins of the oracle's EMFDTD restatement (Yee leapfrog with a complete
CPML and conductive media) against what the mathematics fixes:

 * the CPML tables integrate the auxiliary psi ODE of the convolutional PML
   exactly over one step (checked by numerical integration), with the
   polynomial grading and the outermost layer at index 0;
 * D and H stay divergence-free wherever the plain curl applies (discrete
   Gauss laws: a wrong sign, offset or operand in any curl term breaks them);
 * the hard source pins Ez at the source node to the pulse value;
 * the CPML absorbs: energy leaves the box, and a probe near the PML sees
   what it sees in a box three times larger (the PML disabled fails both);
 * a highly conductive ellipsoid keeps E out of its interior.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

EPS0 = 8.85418781762e-12


def _pml(d, m_a=1.0, m=3.0, k_max=1.0, s_max=0.01222, a_max=1e-3):
    """sigma, kappa, alpha of layer w (0 = innermost PML layer) by the grading."""
    out = []
    for w in range(d):
        wd = (w / (d - 1.0)) ** m
        out.append((wd * s_max, 1.0 + (k_max - 1.0) * wd, ((d - 1.0 - w) / (d - 1.0)) ** m_a * a_max))
    return out


@pytest.mark.parametrize("d,k_max", [(5, 1.0), (10, 1.0), (8, 3.0)])
def test_cpml_tables_integrate_the_psi_ode(d, k_max):
    from scipy.integrate import solve_ivp
    dt = O.emfdtd_dt(0.5)
    b, c, ook = O.cpml_tables(d, k_max=k_max, dt=dt)
    # interior entry: no psi, plain derivative
    assert b[d] == 0.0 and c[d] == 0.0 and ook[d] == 1.0
    for w, (sig, kap, alp) in enumerate(_pml(d, k_max=k_max)):
        idx = d - 1 - w  # reversed: index 0 is the outermost layer
        assert ook[idx] == pytest.approx(1.0 / kap, rel=1e-15)
        # d psi/dt = -((sigma/kappa + alpha)/eps0) psi - (sigma/(eps0 kappa^2)) dE,
        # dE held at 1 over one step, psi(0) = 0.3
        a = (sig / kap + alp) / EPS0
        g = sig / (EPS0 * kap * kap)
        sol = solve_ivp(lambda t, y: -a * y - g, (0.0, dt), [0.3], rtol=1e-12, atol=1e-15, method="DOP853")
        want = sol.y[0, -1]
        got = b[idx] * 0.3 + c[idx] * 1.0
        assert got == pytest.approx(want, rel=1e-9, abs=1e-13), (w, got, want)
        assert 0.0 < b[idx] <= 1.0 and c[idx] <= 0.0


def test_cpml_tables_without_losses_are_identity():
    b, c, ook = O.cpml_tables(6, s_max=0.0, a_max=0.0)
    assert np.all(b[:6] == 1.0) and np.all(c == 0.0) and np.all(ook == 1.0)


def test_ga_gb():
    dt = O.emfdtd_dt(0.5)
    sig = np.array([0.0, 1e-12, 1.0, 3.546e7])
    ga, gb = O.ga_gb(sig, dt)
    assert ga[0] == 1.0 and gb[0] == 0.0
    np.testing.assert_allclose(gb, sig * dt / EPS0, rtol=1e-15)
    np.testing.assert_allclose(ga * (1.0 + gb), 1.0, rtol=1e-15)


def test_source_and_dt_closed_forms():
    dt = O.emfdtd_dt(0.5)
    assert dt == pytest.approx(0.5 / 299792458.0, rel=1e-16)
    for t in (0.0, 17.0, 40.0, 73.0):
        assert O.mod_gauss(t, dt, 0.0, 40.0, 8.0) == pytest.approx(math.exp(-0.5 * ((40.0 - t) / 8.0) ** 2), rel=1e-15)
        want = math.exp(-0.5 * ((40.0 - t) / 8.0) ** 2) * math.cos(2 * math.pi * 3e7 * t * dt)
        assert O.mod_gauss(t, dt, 3e7, 40.0, 8.0) == pytest.approx(want, rel=1e-14, abs=1e-300)


def test_ellipsoid_indicator_on_spheres():
    """The reference's indicator takes theta as the polar angle but puts
    cos(theta) on x, y and sin(theta) on z, so its surface is the true
    ellipsoid only for equal radii (the reference's configurations all use
    spheres).  For a sphere: `inside` within radius - 0.2, ~0 beyond + 0.2."""
    c, r = (10.0, 11.0, 12.0), 5.0
    rng = np.random.default_rng(3)
    inside = outside = 0
    for _ in range(400):
        p = np.array(c) + rng.uniform(-8, 8, 3)
        rho = float(np.linalg.norm(p - c))
        v = O.lib().oracle_ellipsoid(*p, *c, r, r, r, 3.546e7)
        if rho < r - 0.2:
            assert v == pytest.approx(3.546e7, rel=1e-9)
            inside += 1
        elif rho > r + 0.2:
            assert abs(v) < 1e-3
            outside += 1
    assert inside > 20 and outside > 20


def _div_h(S):
    """Discrete div of H at cell centres (i+1/2, j+1/2, k+1/2); arrays [k, j, i]."""
    Hx, Hy, Hz = S["Hx"], S["Hy"], S["Hz"]
    return ((Hx[:-1, :-1, 1:] - Hx[:-1, :-1, :-1]) + (Hy[:-1, 1:, :-1] - Hy[:-1, :-1, :-1]) +
            (Hz[1:, :-1, :-1] - Hz[:-1, :-1, :-1]))


def _div_d(S):
    """Discrete div of D at nodes (i, j, k), i, j, k >= 1."""
    Dx, Dy, Dz = S["Dx"], S["Dy"], S["Dz"]
    return ((Dx[1:, 1:, 1:] - Dx[1:, 1:, :-1]) + (Dy[1:, 1:, 1:] - Dy[1:, :-1, 1:]) +
            (Dz[1:, 1:, 1:] - Dz[:-1, 1:, 1:]))


def test_gauss_laws_hold_where_the_curl_is_plain():
    n, d = 26, 5
    S = O.Emfdtd(n, d=d, src=(13, 12, 14), sink=(13, 12, 18))
    S.run(70)
    scale = max(np.abs(S[f]).max() for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    assert scale > 1e-3  # the pulse is in the box
    lo, hi = d + 1, n - d - 3  # cells whose H / D nodes all sit outside the PML and in range
    dh = _div_h(S)[lo:hi, lo:hi, lo:hi]
    dd = _div_d(S)[lo - 1:hi, lo - 1:hi, lo - 1:hi]
    # round-off of the O(1) fields near the source during the pulse (peak 1)
    assert np.abs(dh).max() < 1e-14
    assert np.abs(dd).max() < 1e-14
    # inside the PML the psi terms make the divergence nonzero: the check is not vacuous
    assert max(np.abs(_div_h(S)).max(), np.abs(_div_d(S)).max()) > 1e-9


def test_hard_source_and_sink():
    S = O.Emfdtd(16, d=4, src=(8, 8, 8), sink=(8, 8, 11))
    for s in range(1, 30):
        v = O.mod_gauss(float(s - 1), S.dt)
        sink = S.step(v)
        assert S["Ez"][8, 8, 8] == v
        assert sink == S["Ez"][11, 8, 8]


def _energy(S):
    return sum(float(np.sum(S[f] ** 2)) for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))


@pytest.mark.parametrize("absorbing", [True, False])
def test_cpml_absorbs_the_pulse(absorbing):
    n = 40
    pml = {} if absorbing else {"s_max": 0.0, "a_max": 0.0}
    S = O.Emfdtd(n, d=8, src=(20, 20, 20), sink=(20, 20, 26), **pml)
    w = []
    for s in range(1, 361):
        S.step(O.mod_gauss(float(s - 1), S.dt))
        if s % 20 == 0:
            w.append(_energy(S))
    # after the pulse (step 80) the radiated energy is in the box; a closed
    # lossless box keeps it, the CPML removes it (the hard source's reactive
    # near field around step 40 is far larger and not a reference)
    ratio = w[-1] / w[3]
    if absorbing:
        assert ratio < 1e-5
    else:
        assert ratio > 0.5


@pytest.fixture(scope="module")
def large_box_probe():
    """Probe 6 cells from the source in a 90^3 box: its own boundary echo
    cannot reach the probe within 150 steps (half a cell per step)."""
    return O.Emfdtd(90, d=8, src=(45, 45, 45), sink=(45, 45, 51)).run(150)


def _tables_variant(kind, d):
    b, c, ook = O.cpml_tables(d)
    if kind == "unreversed":  # sigma_max next to the interior instead of the wall
        return tuple(np.r_[x[:d][::-1], x[d]] for x in (b, c, ook))
    if kind == "c_sign":
        return b, -c, ook
    if kind == "lossless":
        return O.cpml_tables(d, s_max=0.0, a_max=0.0)
    return b, c, ook


@pytest.mark.parametrize("kind", ["right", "unreversed", "c_sign", "lossless"])
def test_cpml_probe_matches_a_larger_box(large_box_probe, kind):
    """The same probe in a 48^3 box with a 12-layer CPML sees the large box's
    signal to 3 % of its peak (the unmodulated pulse is mostly below the CFS
    cutoff alpha/eps0 of the reference's parameters, so the absorption is
    moderate).  The planted table mistakes fail: grading not reversed and
    the psi coefficient's sign flipped by an order of magnitude, a lossless
    layer (a closed box) by more than 1.5x the bound."""
    d, n = 12, 48
    S = O.Emfdtd(n, d=d, src=(24, 24, 24), sink=(24, 24, 30), tables=_tables_variant(kind, d))
    with np.errstate(all="ignore"):
        a = S.run(150)
    b = large_box_probe
    err = np.abs(a - b).max() / np.abs(b).max()
    if kind == "right":
        assert err < 3e-2
    elif kind == "lossless":  # the closed box's echo arrives late in the window
        assert err > 5e-2
    else:
        assert not np.isfinite(err) or err > 0.2


def test_conductor_keeps_e_out():
    n = 30
    sigma = O.ellipsoid_sigma(n, (18, 15, 15), (5, 5, 5))
    S = O.Emfdtd(n, d=5, src=(8, 15, 15), sink=(11, 15, 15), sigma=sigma)
    S.run(120)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    deep = (i - 18) ** 2 + (j - 15) ** 2 + (k - 15) ** 2 < 3.5 ** 2
    e = np.sqrt(S["Ex"] ** 2 + S["Ey"] ** 2 + S["Ez"] ** 2)
    assert e.max() > 1e-3
    assert e[deep].max() < 1e-6 * e.max()
