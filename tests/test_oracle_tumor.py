"""Pins of the oracle's TumorHeatEqnFDM restatement (Pennes bioheat with a
spherical tumor, Neumann flux borders and a Cauchy border on the far face
of the last axis) by what the discrete equations fix:

 * 1-D: the linear profile T(x) = T_inf + q/H + (q/K)(L - x) satisfies both
   border rules exactly (zero-flux balance at x = 0 with inflow q, Cauchy
   outflow at x = L), so without perfusion and source it is a fixed point;
 * 2-D / 3-D: with trapezoid weights (1/2 per face index) the one-sided
   border rules make the scheme conservative: the weighted sum of T changes
   per step by tau/(rho c) times the border fluxes, q on every Neumann face
   and -H (T - T_inf) on the Cauchy face;
 * the tumor selection: omega and Q switch at the sphere's radius.
A wrong factor, sign or face in a border rule fails the first two.
"""
import numpy as np
import pytest

import oracle_lib as O


def _grid(n, x1):
    return O.make_grid(n, x0=[0.0] * len(n), x1=list(x1))


def test_linear_profile_is_a_fixed_point_1d():
    n, L = 41, 0.06
    g = _grid((n,), (L,))
    K, H, q, tinf = 0.6, 10.0, 150.0, 22.4
    h = L / (n - 1)
    tau = 0.4 / (2.0 * K / (1040.0 * 3650.0) / h ** 2 + 2.0 * H / (1040.0 * 3650.0 * h))
    c = O.tumor_coef(tau, (L / 2,), 0.0, k=K, h=H, q_bc=q, t_inf=tinf, omega_brain=0.0, omega_tumor=0.0,
                     q_brain=0.0, q_tumor=0.0)
    x = np.linspace(0.0, L, n)
    prof = tinf + q / H + (q / K) * (L - x)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = prof
    for _ in range(300):
        u = O.tumor_step(g, u, c)
    np.testing.assert_allclose(O.nodes_view(g, u).ravel(), prof, rtol=1e-12, atol=0)
    # a different profile is not a fixed point: the check has teeth
    O.nodes_view(g, u)[...] = prof + 0.5
    u2 = O.tumor_step(g, u, c)
    assert np.abs(O.nodes_view(g, u2) - O.nodes_view(g, u)).max() > 1e-6


@pytest.mark.parametrize("n,x1", [((13, 11), (0.06, 0.03)), ((11, 9, 8), (0.06, 0.03, 0.03))])
def test_weighted_sum_changes_by_the_border_fluxes(n, x1):
    dim = len(n)
    g = _grid(n, x1)
    K, H, q, tinf, rho, cc = 0.6, 10.0, 80.0, 22.4, 1040.0, 3650.0
    h = [x1[p] / (n[p] - 1) for p in range(dim)]
    tau = 0.3 / sum(2.0 * K / (rho * cc) / hp ** 2 for hp in h)
    c = O.tumor_coef(tau, [0.0] * dim, 0.0, k=K, h=H, q_bc=q, t_inf=tinf, omega_brain=0.0, omega_tumor=0.0,
                     q_brain=0.0, q_tumor=0.0)
    rng = np.random.default_rng(3)
    u = np.zeros(O.memall_shape(g))
    T = 30.0 + 5.0 * rng.random(tuple(reversed(n)))
    O.nodes_view(g, u)[...] = T
    un = O.tumor_step(g, u, c)
    dT = O.nodes_view(g, un) - T  # shape (z, y, x) reversed order
    w = np.ones(tuple(reversed(n)))
    for p in range(dim):
        ax = dim - 1 - p
        sl0 = [slice(None)] * dim
        sl1 = [slice(None)] * dim
        sl0[ax], sl1[ax] = 0, n[p] - 1
        w[tuple(sl0)] *= 0.5
        w[tuple(sl1)] *= 0.5
    lhs = float(np.sum(w * dT))
    rhs = 0.0
    for p in range(dim):
        ax = dim - 1 - p
        wt = np.ones(tuple(reversed(n)))
        for r in range(dim):  # transverse trapezoid weights
            if r == p:
                continue
            a = dim - 1 - r
            s0 = [slice(None)] * dim
            s1 = [slice(None)] * dim
            s0[a], s1[a] = 0, n[r] - 1
            wt[tuple(s0)] *= 0.5
            wt[tuple(s1)] *= 0.5
        near = [slice(None)] * dim
        far = [slice(None)] * dim
        near[ax], far[ax] = 0, n[p] - 1
        flux_near = q * wt[tuple(near)]
        if p == dim - 1:
            flux_far = -H * (T[tuple(far)] - tinf) * wt[tuple(far)]
        else:
            flux_far = q * wt[tuple(far)]
        rhs += tau / (rho * cc) * (np.sum(flux_near) + np.sum(flux_far)) / h[p]
    assert lhs == pytest.approx(rhs, rel=1e-10)
    assert abs(rhs) > 1e-6


def test_tumor_selection():
    n = (21, 15, 15)
    g = _grid(n, (0.06, 0.03, 0.03))
    c = O.tumor_coef(1.0, (0.03, 0.015, 0.022), 0.005)
    w, q = O.tumor_fields(g, c)
    x = [np.linspace(0.0, e, m) for e, m in zip((0.06, 0.03, 0.03), n)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    d2 = (X - 0.03) ** 2 + (Y - 0.015) ** 2 + (Z - 0.022) ** 2
    inside = d2 < 0.005 ** 2 * (1 - 1e-9)
    outside = d2 > 0.005 ** 2 * (1 + 1e-9)
    wn = O.nodes_view(g, w)
    assert np.all(wn[inside] == 0.0007) and np.all(wn[outside] == 0.004) and inside.sum() > 5
