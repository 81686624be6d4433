"""Pins of the oracle's Pennes step with the Tumor example's border rule
(TumorHeatEqnFDM::updateBorder, examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:272-349)
on node fields, of its adjoint step and of the perfusion gradient:

 * the step on the fields of oracle_tumor_fields equals oracle_tumor_step
   (the restated example, pinned by its own closed forms) bit for bit;
 * the adjoint step equals A^T lam for the matrix A assembled column by
   column from the step's linear part (unit vectors, zero source, ambient,
   flux and arterial temperature) on tiny 1-D, 2-D and 3-D grids, ghost
   layers filled with NaN so that no read beyond the grid goes unnoticed;
   a transpose that forgets the doubled face weights fails it;
 * dJ/dw against central finite differences of the oracle's forward, along
   a random direction and per node on faces, edges, corners and inside.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

EXTENT = (0.06, 0.03, 0.03)


def _case(n, nlayers=1, q_bc=120.0):
    dim = len(n)
    g = O.make_grid(n, x0=[0.0] * dim, x1=list(EXTENT[:dim]), nlayers=nlayers)
    h = [O.lib().oracle_gridsize(C.byref(g), p) for p in range(dim)]
    rho, cc, K, H = 1040.0, 3650.0, 0.6, 10.0
    tau = 0.5 / (sum(2.0 * K / (rho * cc) / hp ** 2 for hp in h) + 2.0 * H / (rho * cc * h[-1]) +
                 1052.0 * 3800.0 / (rho * cc) * 0.004)
    center = [e / 2.0 for e in EXTENT[:dim - 1]] + [EXTENT[dim - 1] - 0.008]
    tc = O.tumor_coef(tau, center, 0.005, q_bc=q_bc)
    coef = O.pennes_coef(tau=tau, cdiff=tc.k / (tc.rho * tc.c), cperf=(tc.rho_b * tc.c_pb) / (tc.rho * tc.c),
                         csrc=1.0 / (tc.rho * tc.c), t_a=tc.t_a)
    border = O.tumor_border(tc.k, tc.h, tc.t_inf, tc.q_bc)
    return g, tc, coef, border


def _domain(g):
    m = np.zeros(O.memall_shape(g), bool)
    O.nodes_view(g, m)[...] = True
    return m


@pytest.mark.parametrize("n", [(31,), (25, 13), (21, 11, 12)])
@pytest.mark.parametrize("nlayers", [1, 2])
def test_step_equals_restated_example(n, nlayers):
    g, tc, coef, border = _case(n, nlayers)
    w, q = O.tumor_fields(g, tc)
    rng = np.random.default_rng(3)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = 37.0 + rng.standard_normal(tuple(reversed(n)))
    a = b = u
    for _ in range(10):
        a = O.tumor_step(g, a, tc)
        b = O.pennes_tumor_step(g, b, w, q, coef, border)
    np.testing.assert_array_equal(O.nodes_view(g, a), O.nodes_view(g, b))


def _linear_case(n):
    g, tc, coef, border = _case(n)
    lin = O.pennes_coef(tau=coef.tau, cdiff=coef.cdiff, cperf=coef.cperf, csrc=coef.csrc, t_a=0.0)
    lb = O.tumor_border(border.k, border.h, 0.0, 0.0)
    rng = np.random.default_rng(5)
    w = 0.004 * np.exp(0.5 * rng.standard_normal(O.memall_shape(g)))
    return g, lin, lb, w


def _matrix(g, lin, lb, w):
    dom = _domain(g)
    idx = np.flatnonzero(dom.ravel())
    zero = np.zeros(O.memall_shape(g))
    A = np.empty((idx.size, idx.size))
    for c, m in enumerate(idx):
        e = np.full(O.memall_shape(g), np.nan)
        e[dom] = 0.0
        e.ravel()[m] = 1.0
        A[:, c] = O.pennes_tumor_step(g, e, w, zero, lin, lb).ravel()[idx]
    return A, dom, idx


def _naive_transpose(g, lam, w, lin):
    """The Keep-border adjoint's central form on every node (no doubled face
    weights, no Cauchy diagonal): a plausible mistake the pin must reject."""
    out = lam.copy()
    st = [1, O.memall_shape(g)[2], O.memall_shape(g)[2] * O.memall_shape(g)[1]]
    flat, res = np.nan_to_num(lam).ravel(), out.ravel()
    dom = _domain(g).ravel()
    for m in np.flatnonzero(dom):
        v = flat[m]
        for p in range(g.dim):
            hp = O.lib().oracle_gridsize(C.byref(g), p)
            v += lin.tau * lin.cdiff * (flat[m - st[p]] + flat[m + st[p]] - 2 * flat[m]) / hp ** 2
        res[m] = v - lin.tau * lin.cperf * w.ravel()[m] * flat[m]
    return out


@pytest.mark.parametrize("n", [(6,), (5, 4), (4, 5, 4), (3, 3, 3)])
def test_adjoint_step_is_the_transpose(n):
    g, lin, lb, w = _linear_case(n)
    A, dom, idx = _matrix(g, lin, lb, w)
    rng = np.random.default_rng(9)
    lam = np.full(O.memall_shape(g), np.nan)
    lam[dom] = rng.standard_normal(idx.size)
    got = O.pennes_tumor_adjoint_step(g, lam, w, lin, lb).ravel()[idx]
    want = A.T @ lam.ravel()[idx]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    bad = _naive_transpose(g, lam, w, lin).ravel()[idx]
    assert np.abs(bad - want).max() > 1e-3 * np.abs(want).max()


def test_adjoint_dot_product_3d():
    g, lin, lb, w = _linear_case((13, 9, 8))
    dom = _domain(g)
    rng = np.random.default_rng(2)
    d = np.full(O.memall_shape(g), np.nan)
    lam = np.full(O.memall_shape(g), np.nan)
    d[dom] = rng.standard_normal(dom.sum())
    lam[dom] = rng.standard_normal(dom.sum())
    zero = np.zeros(O.memall_shape(g))
    Ad = O.pennes_tumor_step(g, d, w, zero, lin, lb)
    Atl = O.pennes_tumor_adjoint_step(g, lam, w, lin, lb)
    lhs, rhs = np.dot(lam[dom], Ad[dom]), np.dot(Atl[dom], d[dom])
    assert abs(lhs - rhs) <= 1e-12 * np.abs(lam[dom] * Ad[dom]).sum()


def _grad_case(n=(9, 7, 6), nsteps=8):
    g, tc, coef, border = _case(n, q_bc=120.0)
    w, q = O.tumor_fields(g, tc)
    rng = np.random.default_rng(17)
    t0 = np.zeros(O.memall_shape(g))
    O.nodes_view(g, t0)[...] = 37.0 + 0.5 * rng.standard_normal(tuple(reversed(n)))
    w_t = w * (1.0 + 0.3 * rng.standard_normal(w.shape))
    target = O.pennes_tumor_forward_all(g, t0, w_t, q, coef, border, nsteps)[-1]
    return g, coef, border, w, q, t0, target, nsteps


def _J(g, coef, border, w, q, t0, target, nsteps):
    tN = O.pennes_tumor_forward_all(g, t0, w, q, coef, border, nsteps)[-1]
    return 0.5 * np.sum((O.nodes_view(g, tN) - O.nodes_view(g, target)) ** 2)


def test_gradient_vs_finite_differences():
    g, coef, border, w, q, t0, target, N = _grad_case()
    st = O.pennes_tumor_forward_all(g, t0, w, q, coef, border, N)
    J, grad = O.pennes_tumor_grad(g, st, w, target, coef, border, N)
    assert J == pytest.approx(_J(g, coef, border, w, q, t0, target, N), rel=1e-14)
    dom = _domain(g)
    assert np.all(grad[~dom] == 0.0)
    rng = np.random.default_rng(4)
    v = np.where(dom, w * rng.standard_normal(w.shape), 0.0)
    eps = 1e-5  # central differences: error ~ eps^2 (2.7e-6 relative at 1e-3, 4.6e-9 at 1e-5)
    fd = (_J(g, coef, border, w + eps * v, q, t0, target, N) -
          _J(g, coef, border, w - eps * v, q, t0, target, N)) / (2 * eps)
    assert np.dot(grad.ravel(), v.ravel()) == pytest.approx(fd, rel=1e-7)
    # per node: corners (one on the Cauchy face), face centres, an inner node
    n = [g.n[p] for p in range(3)]
    nodes = [(0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), (n[0] // 2, n[1] // 2, n[2] - 1),
             (0, n[1] // 2, n[2] // 2), (n[0] - 1, 1, n[2] // 2), (n[0] // 2, 0, 0), (n[0] // 2, n[1] // 2, 2)]
    for (i, j, k) in nodes:
        e = np.zeros(w.shape)
        O.nodes_view(g, e)[k, j, i] = 1.0
        # truncation ~ h^2 (1.6e-5 relative at the corner for 1e-3 w_max),
        # rounding of T ~ 37 over 2h (about 1e-7 absolute at 1e-5 w_max)
        h = 1e-4 * w.max()
        fd = (_J(g, coef, border, w + h * e, q, t0, target, N) -
              _J(g, coef, border, w - h * e, q, t0, target, N)) / (2 * h)
        gi = O.nodes_view(g, grad)[k, j, i]
        assert gi == pytest.approx(fd, rel=2e-6, abs=2e-8), (i, j, k)
