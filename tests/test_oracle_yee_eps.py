"""Pins of the oracle's Yee step with a relative permittivity field eps and
of its discrete adjoint gradient dJ/deps (J = 0.5 sum over updated E entries
of (E^N - target)^2):

 * eps = 1 everywhere: bit-identical to the vacuum step;
 * the eps-weighted discrete energy 0.5 sum eps E.E + 0.5 sum H^{n-1/2}.H^{n+1/2}
   is conserved to rounding (it is only for E += (s/eps(m)) curl H with the
   node's own eps; a neighbour's eps or eps instead of 1/eps breaks it);
 * the gradient against central finite differences of J, every component on
   a tiny grid (brute force) and a random direction on a larger one;
 * constant eps = a: the summed field gradient equals d/da of the scalar
   problem;
 * sensitivity: a vectorised numpy restatement agrees with the oracle to
   rounding, and the same restatement with a planted mistake (sign of the
   H-adjoint term, lam_E left nonzero on PEC-tangential entries, the -x
   neighbour's eps in the gradient factor) fails the finite differences.
"""
import numpy as np
import pytest

import oracle_lib as O


def _case(N, seed=0, full=False):
    rng = np.random.default_rng(seed)
    m = N + 1
    f0 = rng.standard_normal((6, m, m, m))
    if not full:
        f0[:3] *= O.yee_emask(N)
    eps = 1.0 + rng.random((m, m, m))
    target = 0.3 * rng.standard_normal((3, m, m, m))
    return f0, eps, target


def _J(N, s, eps, f0, target, nsteps):
    f = O.yee_eps_step(N, s, eps, f0.copy(), nsteps)
    r = np.where(O.yee_emask(N), f[:3] - target, 0.0)
    return 0.5 * float(np.sum(r * r))


def test_unit_eps_is_the_vacuum_step():
    N, s = 9, 0.5
    f0, _, _ = _case(N, seed=1, full=True)
    a = O.yee_step(N, s, f0.copy(), 7)
    b = O.yee_eps_step(N, s, np.ones((N + 1,) * 3), f0.copy(), 7)
    np.testing.assert_array_equal(a, b)


def _weighted_energy(N, f, h_prev, eps):
    return 0.5 * float(np.sum(eps[None] * f[:3] * f[:3])) + 0.5 * float(np.sum(f[3:] * h_prev))


@pytest.mark.parametrize("variant", ["own", "neighbour", "times_eps"])
def test_weighted_energy_invariant(variant):
    N, s = 10, 0.5
    f, eps, _ = _case(N, seed=2)
    if variant == "own":
        step_e = lambda f: O.lib().oracle_yee_eps_half_e(N, s, O.ptr(eps), O.ptr(f))
    else:
        # planted mistakes, restated in numpy: the -x neighbour's eps, or eps
        # in place of 1/eps
        e_used = np.roll(eps, 1, axis=2) if variant == "neighbour" else 1.0 / eps

        def step_e(f):
            before = f[:3].copy()
            O.yee_half_e(N, s, f)
            f[:3] = before + (f[:3] - before) / e_used[None]
    W = []
    for _ in range(40):
        h_prev = f[3:].copy()
        O.yee_half_h(N, s, f)
        W.append(_weighted_energy(N, f, h_prev, eps))
        step_e(f)
    W = np.array(W)
    drift = np.abs(W - W[0]).max() / abs(W[0])
    if variant == "own":
        assert drift < 1e-12
    else:
        assert drift > 1e-3


def test_gradient_every_component_vs_finite_differences():
    N, s, nsteps = 4, 0.5, 6
    f0, eps, target = _case(N, seed=3, full=True)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    J, g = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    assert J == pytest.approx(_J(N, s, eps, f0, target, nsteps), rel=1e-14)
    h = 1e-6
    fd = np.zeros(eps.shape)
    for idx in np.ndindex(eps.shape):
        ep, em = eps.copy(), eps.copy()
        ep[idx] += h
        em[idx] -= h
        fd[idx] = (_J(N, s, ep, f0, target, nsteps) - _J(N, s, em, f0, target, nsteps)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())
    # nodes where no E component is updated (the corner lines) carry no gradient
    anym = O.yee_emask(N).any(axis=0)
    assert np.all(g[~anym] == 0.0) and np.count_nonzero(np.abs(fd[anym]) > 1e-9) == int(anym.sum())


def test_gradient_random_direction():
    N, s, nsteps = 12, 0.55, 15
    f0, eps, target = _case(N, seed=4)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    J, g = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    v = np.random.default_rng(5).standard_normal(eps.shape)
    h = 1e-6
    fd = (_J(N, s, eps + h * v, f0, target, nsteps) - _J(N, s, eps - h * v, f0, target, nsteps)) / (2 * h)
    assert float(np.sum(g * v)) == pytest.approx(fd, rel=1e-7)


def test_constant_eps_scalar_derivative():
    N, s, nsteps = 8, 0.5, 10
    f0, eps, target = _case(N, seed=6)
    eps[...] = 1.7
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    _, g = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    h = 1e-6
    fd = (_J(N, s, eps + h, f0, target, nsteps) - _J(N, s, eps - h, f0, target, nsteps)) / (2 * h)
    assert float(np.sum(g)) == pytest.approx(fd, rel=1e-7)


# -- vectorised numpy restatement (arrays [k, j, i]) with switches ----------
def _fwd(a, ax):
    return np.roll(a, -1, ax) - a


def _bwd(a, ax):
    return a - np.roll(a, 1, ax)


def _hform(x, y, z):
    return np.stack([_fwd(y, 0) - _fwd(z, 1), _fwd(z, 2) - _fwd(x, 0), _fwd(x, 1) - _fwd(y, 2)])


def _eform(x, y, z):
    return np.stack([_bwd(z, 1) - _bwd(y, 0), _bwd(x, 0) - _bwd(z, 2), _bwd(y, 2) - _bwd(x, 1)])


def _hmask(N):
    m = N + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    return np.stack([(j < N) & (k < N), (i < N) & (k < N), (i < N) & (j < N)])


def _np_grad(N, s, eps, st, target, nsteps, flip=False, keep_tangential=False, nb_eps=False):
    em, hm = O.yee_emask(N), _hmask(N)
    ce = s / eps
    lamE = np.where(em, st[nsteps, :3] - target, 0.0)
    lamH = np.zeros_like(lamE)
    grad = np.zeros(eps.shape)
    for n in range(nsteps, 0, -1):
        E, H = st[n - 1, :3], st[n - 1, 3:]
        P = np.where(hm, H + s * _hform(*E), H)
        mu = np.where(hm, lamH + (1.0 if flip else -1.0) * _hform(*(ce[None] * lamE)), lamH)
        acc = np.sum(np.where(em, lamE * _eform(*P), 0.0), axis=0)
        e_g = np.roll(eps, 1, axis=2) if nb_eps else eps
        grad -= np.where(em.any(axis=0), (ce / e_g) * acc, 0.0)
        new = lamE - s * _eform(*mu)
        lamE = new if keep_tangential else np.where(em, new, 0.0)
        lamH = mu
    return grad


def test_numpy_restatement_and_planted_mistakes():
    N, s, nsteps = 5, 0.5, 6
    f0, eps, target = _case(N, seed=8)
    st = O.yee_eps_forward_all(N, s, eps, f0, nsteps)
    _, g = O.yee_eps_grad(N, s, eps, st, target, nsteps)
    good = _np_grad(N, s, eps, st, target, nsteps)
    np.testing.assert_allclose(good, g, rtol=1e-10, atol=1e-13 * np.abs(g).max())
    v = np.random.default_rng(9).standard_normal(eps.shape)
    h = 1e-6
    fd = (_J(N, s, eps + h * v, f0, target, nsteps) - _J(N, s, eps - h * v, f0, target, nsteps)) / (2 * h)
    assert float(np.sum(good * v)) == pytest.approx(fd, rel=1e-7)
    for kw in ({"flip": True}, {"keep_tangential": True}, {"nb_eps": True}):
        bad = _np_grad(N, s, eps, st, target, nsteps, **kw)
        assert abs(float(np.sum(bad * v)) - fd) > 1e-4 * abs(fd), kw
