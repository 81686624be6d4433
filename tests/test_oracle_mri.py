"""Pins of the oracle's MRIDataPBHEqnFDM restatement (Pennes bioheat with
per-node material fields, Neumann borders, open-skull convection and
radiation or closed-skull flux on the far face of the last axis):

 * with uniform materials, no radiation and an open skull everywhere it is
   the Tumor example's update, bit for bit (an independently pinned oracle);
 * 1-D fixed points: the linear profile with inflow q_bc at x = 0 and
   convection + radiation (open skull) or flux -q_bc (closed skull) at x = L;
 * the discrete balance of the non-conservative form with node materials:
   sum_i w_i (rho c / lambda)_i dT_i / tau = sum over border faces of the
   face fluxes / (lambda_i h_p), trapezoid weights w_i;
 * the radiation term evaluates T^4 at the value accumulated in the new
   field before the last axis (the reference reads vNew there).
"""
import numpy as np
import pytest

import oracle_lib as O

EXT = (0.06, 0.03, 0.03)


def _grid(n):
    return O.make_grid(n, x0=[0.0] * len(n), x1=list(EXT[:len(n)]))


def _mats(n, rng=None, surface=None, uniform=None):
    shape = tuple(reversed(n))
    if uniform is not None:
        vals = [np.full(shape, v) for v in uniform]
    else:
        vals = [1040.0 * (1 + 0.2 * rng.random(shape)), 3650.0 * (1 + 0.2 * rng.random(shape)),
                0.5 + 0.2 * rng.random(shape), np.full(shape, 1052.0), np.full(shape, 3800.0),
                0.004 * rng.random(shape), 37.0 + rng.random(shape), 25000.0 * rng.random(shape),
                np.zeros(shape)]
    if surface is not None:
        vals[8] = surface.astype(np.float64)
    return vals


def _tau(n, lam=0.7, rho=1040.0, c=3650.0):
    h = [EXT[p] / (n[p] - 1) for p in range(len(n))]
    return 0.3 / sum(2.0 * lam / (rho * c) / hp ** 2 for hp in h)


@pytest.mark.parametrize("n", [(15,), (13, 9), (11, 9, 8)])
def test_equals_tumor_update_with_uniform_materials(n):
    g = _grid(n)
    tau = _tau(n)
    center = [e / 2 for e in EXT[:len(n) - 1]] + [EXT[len(n) - 1] - 0.008]
    tc = O.tumor_coef(tau, center, 0.006, q_bc=50.0, h=10.0, t_inf=22.4)
    w, q = O.tumor_fields(g, tc)
    uni = [tc.rho, tc.c, tc.k, tc.rho_b, tc.c_pb, 0.0, tc.t_a, 0.0, 1.0]
    mats = _mats(n, uniform=uni)
    mats[5] = O.nodes_view(g, w).copy()
    mats[7] = O.nodes_view(g, q).copy()
    mc = O.mri_coef(tau, h=tc.h, t_amb=tc.t_inf, q_bc=tc.q_bc, epsilon=0.0)
    rng = np.random.default_rng(2)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = 36.0 + rng.random(tuple(reversed(n)))
    np.testing.assert_array_equal(O.mri_step(g, u, mats, mc), O.tumor_step(g, u, tc))


def _radiative_tl(qbc, H, tamb, eps, sig):
    """T_L with q_bc = H (T_L - T_amb) + eps sig ((T_L+273.15)^4 - (T_amb+273.15)^4)."""
    t = tamb + 1.0
    for _ in range(60):
        f = H * (t - tamb) + eps * sig * ((t + 273.15) ** 4 - (tamb + 273.15) ** 4) - qbc
        df = H + 4 * eps * sig * (t + 273.15) ** 3
        t -= f / df
    return t


@pytest.mark.parametrize("skull", ["open", "closed"])
def test_1d_fixed_points(skull):
    n = 41
    g = _grid((n,))
    lam, qbc, H, tamb, eps, sig = 0.55, 120.0, 10.0, 22.0, 0.98, 5.670373e-8
    tau = _tau((n,), lam)
    x = np.linspace(0.0, EXT[0], n)
    if skull == "open":
        tl = _radiative_tl(qbc, H, tamb, eps, sig)
        mc = O.mri_coef(tau, h=H, t_amb=tamb, q_bc=qbc, epsilon=eps, sigma=sig)
        surface = np.ones(n)
    else:
        tl = 30.0
        mc = O.mri_coef(tau, h=H, t_amb=tamb, q_bc=qbc, q_skull=-qbc, epsilon=eps, sigma=sig)
        surface = np.zeros(n)
    prof = tl + (qbc / lam) * (EXT[0] - x)
    mats = _mats((n,), uniform=[1040.0, 3650.0, lam, 1052.0, 3800.0, 0.0, 37.0, 0.0, 0.0], surface=surface)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = prof
    for _ in range(200):
        u = O.mri_step(g, u, mats, mc)
    np.testing.assert_allclose(O.nodes_view(g, u).ravel(), prof, rtol=1e-11, atol=0)
    # a shifted profile moves under convection; with flux borders only (closed
    # skull) the level is free but the slope is fixed by q_bc
    O.nodes_view(g, u)[...] = prof + 0.5 if skull == "open" else tl + 1.1 * (qbc / lam) * (EXT[0] - x)
    assert np.abs(O.nodes_view(g, O.mri_step(g, u, mats, mc)) - O.nodes_view(g, u)).max() > 1e-7


def _weights(n):
    w = np.ones(tuple(reversed(n)))
    dim = len(n)
    for p in range(dim):
        ax = dim - 1 - p
        for e in (0, n[p] - 1):
            sl = [slice(None)] * dim
            sl[ax] = e
            w[tuple(sl)] *= 0.5
    return w


@pytest.mark.parametrize("n,eps", [((11, 9, 8), 0.0), ((13, 9), 0.9)])
def test_discrete_balance_with_node_materials(n, eps):
    dim = len(n)
    g = _grid(n)
    rng = np.random.default_rng(4)
    top = [slice(None)] * dim
    surface = np.zeros(tuple(reversed(n)))
    top[0] = n[dim - 1] - 1
    surface[tuple(top)] = (rng.random(surface[tuple(top)].shape) < 0.5)
    mats = _mats(n, rng=rng, surface=surface)
    mats[5][...] = 0.0  # no perfusion
    mats[7][...] = 0.0  # no source
    tau = _tau(n)
    qbc, qsk, H, tamb, sig = 40.0, -25.0, 10.0, 22.0, 5.670373e-8
    mc = O.mri_coef(tau, h=H, t_amb=tamb, q_bc=qbc, q_skull=qsk, epsilon=eps, sigma=sig)
    T = 30.0 + 5.0 * rng.random(tuple(reversed(n)))
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = T
    dT = O.nodes_view(g, O.mri_step(g, u, mats, mc)) - T
    rho, c, lam = mats[0], mats[1], mats[2]
    lhs = float(np.sum(_weights(n) * rho * c / lam * dT)) / tau
    h = [EXT[p] / (n[p] - 1) for p in range(dim)]
    rhs = 0.0
    for p in range(dim):
        ax = dim - 1 - p
        wt = np.ones(tuple(reversed(n)))
        for r in range(dim):
            if r != p:
                for e in (0, n[r] - 1):
                    sl = [slice(None)] * dim
                    sl[dim - 1 - r] = e
                    wt[tuple(sl)] *= 0.5
        near = [slice(None)] * dim
        far = [slice(None)] * dim
        near[ax], far[ax] = 0, n[p] - 1
        near, far = tuple(near), tuple(far)
        rhs += np.sum(wt[near] * qbc / lam[near]) / h[p]
        if p < dim - 1:
            rhs += np.sum(wt[far] * qbc / lam[far]) / h[p]
        else:
            open_ = surface[far] == 1
            Tf = T[far]
            # radiation at the value accumulated before the last axis
            acc = Tf.copy()
            for r in range(dim - 1):
                ar = dim - 1 - r
                lo = np.take(T, [n[dim - 1] - 1], axis=0).squeeze(0)
                tv = np.take(T, [0], axis=0)  # placeholder to keep shapes explicit
                del tv
                Tm = np.roll(lo, 1, axis=ar - 1)
                Tp = np.roll(lo, -1, axis=ar - 1)
                idx = np.arange(n[r]).reshape([-1 if q == ar - 1 else 1 for q in range(dim - 1)])
                term = np.where(idx == n[r] - 1, (Tm + Tm) - ((2.0 * h[r] / lam[far]) * (-1.0 * qbc)) - 2.0 * lo,
                                np.where(idx == 0, (Tp + ((2.0 * h[r] / lam[far]) * qbc)) + Tp - 2.0 * lo,
                                         (Tm + Tp) - 2.0 * lo))
                acc = acc + tau * (lam[far] / (rho[far] * c[far])) * term / (h[r] * h[r])
            rad = eps * sig * ((acc + 273.15) ** 4 - (tamb + 273.15) ** 4)
            flux = np.where(open_, -H * (Tf - tamb) - rad, qsk)
            rhs += np.sum(wt[far] * flux / lam[far]) / h[p]
    assert lhs == pytest.approx(rhs, rel=1e-9)
    assert abs(rhs) > 1.0
