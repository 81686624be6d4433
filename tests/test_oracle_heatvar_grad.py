"""Pins of the oracle's discrete adjoint of the variable-kappa heat step
w.r.t. the kappa field (J = 0.5 sum_inner (u^N - target)^2):

 * every component against central finite differences of J on a tiny grid
   (brute force, both border modes, one and two border layers);
 * a random direction on a 3-D grid (directional derivative);
 * constant kappa: the field gradient summed over all nodes equals dJ/dkappa
   of a scalar kappa, which the chain rule gives as
   -sum_n tau <lam^n, L u^{n-1}> ... checked here against a finite
   difference of the scalar problem;
 * sensitivity: a planted sign flip of the face term, or lam left nonzero
   on border nodes, breaks the brute-force check.
"""
import numpy as np
import pytest

import oracle_lib as O


def _case(n, nlayers=1, seed=0, nsteps=4, border=O.BORDER_KEEP):
    g = O.make_grid(n, nlayers=nlayers)
    rng = np.random.default_rng(seed)
    shape = O.memall_shape(g)
    u0 = np.zeros(shape)
    O.nodes_view(g, u0)[...] = rng.random(tuple(reversed(n)))
    kap = np.ones(shape)
    O.nodes_view(g, kap)[...] = 0.5 + rng.random(tuple(reversed(n)))
    h = [1.0 / (m - 1) for m in n]
    tau = 0.8 / sum(2.0 * 1.5 / hp ** 2 for hp in h)
    target = np.zeros(shape)
    O.nodes_view(g, target)[...] = 0.5 * rng.random(tuple(reversed(n)))
    return g, u0, kap, tau, target, nsteps, border


def _J(g, u0, kap, tau, target, nsteps, border):
    uN = O.heat7var_forward_all(g, u0, kap, tau, nsteps, border, 0.3)[-1]
    inner = np.zeros(u0.shape, bool)
    sl = tuple(slice(g.hw + g.nlayers, g.hw + g.n[2 - a] - g.nlayers) if g.n[2 - a] > 1 else slice(None)
               for a in range(3))
    inner[sl] = True
    r = np.where(inner, uN - target, 0.0)
    return 0.5 * float(np.sum(r * r))


def _grad(g, u0, kap, tau, target, nsteps, border):
    st = O.heat7var_forward_all(g, u0, kap, tau, nsteps, border, 0.3)
    return O.heat7var_grad(g, st, kap, target, tau, nsteps)


@pytest.mark.parametrize("n,nlayers,border", [((5, 4, 4), 1, O.BORDER_KEEP), ((6, 5), 1, O.BORDER_VALUE),
                                              ((6, 6, 5), 2, O.BORDER_KEEP), ((7,), 1, O.BORDER_VALUE)])
def test_every_component_vs_finite_differences(n, nlayers, border):
    g, u0, kap, tau, target, N, _ = _case(n, nlayers)
    J, grad = _grad(g, u0, kap, tau, target, N, border)
    assert J == pytest.approx(_J(g, u0, kap, tau, target, N, border), rel=1e-14)
    eps = 1e-6
    vshape = O.nodes_view(g, kap).shape
    fd = np.zeros(vshape)
    for idx in np.ndindex(vshape):
        kp, km = kap.copy(), kap.copy()
        O.nodes_view(g, kp)[tuple(idx)] += eps
        O.nodes_view(g, km)[tuple(idx)] -= eps
        fd[tuple(idx)] = (_J(g, u0, kp, tau, target, N, border) - _J(g, u0, km, tau, target, N, border)) / (2 * eps)
    gv = O.nodes_view(g, grad)
    np.testing.assert_allclose(gv, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
    assert np.abs(fd).max() > 1e-4
    n_inner = int(np.prod([max(m - 2 * nlayers, 1) for m in n]))
    assert np.count_nonzero(np.abs(fd) > 1e-9 * np.abs(fd).max()) >= n_inner  # inner nodes and their neighbours


def test_random_direction_3d():
    g, u0, kap, tau, target, N, border = _case((12, 11, 10), seed=3, nsteps=9)
    J, grad = _grad(g, u0, kap, tau, target, N, border)
    rng = np.random.default_rng(7)
    v = np.zeros(kap.shape)
    O.nodes_view(g, v)[...] = rng.standard_normal(tuple(reversed((12, 11, 10))))
    eps = 1e-6
    fd = (_J(g, u0, kap + eps * v, tau, target, N, border) - _J(g, u0, kap - eps * v, tau, target, N, border)) / (2 * eps)
    assert float(np.sum(grad * v)) == pytest.approx(fd, rel=1e-7)


def test_constant_kappa_scalar_derivative():
    """kappa = c everywhere: sum of the field gradient = dJ/dc."""
    g, u0, kap, tau, target, N, border = _case((9, 8, 7), seed=5, nsteps=6)
    O.nodes_view(g, kap)[...] = 0.9
    J, grad = _grad(g, u0, kap, tau, target, N, border)
    eps = 1e-6
    allk = np.ones(kap.shape)
    fd = (_J(g, u0, kap + eps * allk, tau, target, N, border) -
          _J(g, u0, kap - eps * allk, tau, target, N, border)) / (2 * eps)
    assert float(np.sum(O.nodes_view(g, grad))) == pytest.approx(fd, rel=1e-7)


def test_planted_mistakes_break_the_pin():
    """A restatement with the face term's sign flipped, or with lam kept on
    border nodes, fails the finite-difference comparison."""
    n = (5, 4, 4)
    g, u0, kap, tau, target, N, border = _case(n)
    st = O.heat7var_forward_all(g, u0, kap, tau, N, border, 0.3)
    J, grad = O.heat7var_grad(g, st, kap, target, tau, N)
    # numpy restatement of the same adjoint with switches for the mistakes
    hh = [(1.0 / (m - 1)) ** 2 for m in n]

    def adj(flip=False, keep_border=False):
        lam = np.zeros(kap.shape)
        inner = np.zeros(kap.shape, bool)
        inner[2:-2, 2:-2, 2:-2] = True
        lam[inner] = (st[N] - target)[inner]
        gr = np.zeros(kap.shape)
        for s in range(N, 0, -1):
            u = st[s - 1]
            for ax, p in ((2, 0), (1, 1), (0, 2)):
                for sh in (1, -1):
                    nb_l, nb_u = np.roll(lam, sh, ax), np.roll(u, sh, ax)
                    valid = np.zeros(kap.shape, bool)
                    valid[1:-1, 1:-1, 1:-1] = True
                    valid &= np.roll(valid, sh, ax)
                    term = (lam - nb_l) * (nb_u - u) * (-1.0 if flip else 1.0)
                    gr += np.where(valid, 0.5 * (tau / hh[p]) * term, 0.0)
            new = lam.copy()
            for ax, p in ((2, 0), (1, 1), (0, 2)):
                lw, le = np.roll(lam, 1, ax), np.roll(lam, -1, ax)
                kw, ke = np.roll(kap, 1, ax), np.roll(kap, -1, ax)
                new += tau * (0.5 * (ke + kap) * (le - lam) - 0.5 * (kap + kw) * (lam - lw)) / hh[p]
            lam = np.where(inner, new, new if keep_border else 0.0)
            lam[~np.pad(np.ones(tuple(m - 0 for m in reversed(n)), bool), 1)] = 0.0
        return gr
    good = adj()
    np.testing.assert_allclose(O.nodes_view(g, good), O.nodes_view(g, grad), rtol=1e-10, atol=1e-14)
    for bad in (adj(flip=True), adj(keep_border=True)):
        assert np.abs(O.nodes_view(g, bad) - O.nodes_view(g, grad)).max() > 1e-3 * np.abs(grad).max()
