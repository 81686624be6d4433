"""Synthetic workloads of the benchmark configurations (host-side twins of
the device init kernels, used to build identical inputs on the CPU)."""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, idx):
    """splitmix64 of (seed, idx) for a uint64 index array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def batch2d_ensemble(nprob, nsteps, seed=0, sizes=None, nmin=32, nmax=256):
    """Synthetic ensemble of 2-D heat problems (host arrays for scafes_batch2d_run).

    n_i ~ U{nmin..nmax} (seeded), kappa_i ~ U[0.1, 1], tau_i = 0.9 h_i^2 / 4
    (stable for every kappa <= 1 and independent of kappa_i, so dJ/dkappa is
    taken at fixed tau), u0 = sin(pi x) sin(pi y) + 0.01 seeded noise on
    inner nodes, target = u0 * exp(-2 pi^2 * 1.1 kappa_i * nsteps tau_i)
    (the fundamental mode decayed with a 10 % larger diffusivity).
    order: descending per-thread block edge ceil(n/32) (a B x B block)."""
    rng = np.random.default_rng(seed)
    n = np.asarray(sizes if sizes is not None else rng.integers(nmin, nmax + 1, size=nprob), dtype=np.int32)
    nprob = len(n)
    kappa = rng.uniform(0.1, 1.0, size=nprob).astype(np.float32)
    h = 1.0 / (n.astype(np.float64) - 1)
    tau = (0.9 * h * h / 4.0).astype(np.float32)
    off = np.zeros(nprob, dtype=np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    total = int(np.sum(n.astype(np.int64) ** 2))
    u0 = np.zeros(total, dtype=np.float32)
    target = np.zeros(total, dtype=np.float32)
    for i in range(nprob):
        m = int(n[i])
        x = np.linspace(0.0, 1.0, m)
        s = np.sin(np.pi * x)
        u = np.outer(s, s)
        if m > 2:
            u[1:-1, 1:-1] += 0.01 * rng.standard_normal((m - 2, m - 2))
        u[0, :] = u[-1, :] = u[:, 0] = u[:, -1] = 0.0
        dec = np.exp(-2 * np.pi ** 2 * 1.1 * float(kappa[i]) * nsteps * float(tau[i]))
        u0[off[i]:off[i] + m * m] = u.ravel()
        target[off[i]:off[i] + m * m] = (u * dec).ravel()
    cost = (n + 31) // 32
    order = np.argsort(-cost, kind="stable").astype(np.int32)
    return {"n": n, "off": off, "kappa": kappa, "tau": tau, "u0": u0, "target": target, "order": order}


def pennes_workload(n, seed=0, k=0.5, amp=50.0, width=0.1):
    """Pennes bioheat workload on n^3 nodes of the unit cube (memAll arrays
    with one ghost layer, shape (n+2,)*3): log-normal perfusion
    w = 0.5 exp(0.3 Z) (Z i.i.d. N(0,1), seeded), Gaussian source
    Q = amp exp(-|x - c|^2 / (2 width^2)), tau = 0.9 / (6 k / h^2 + max w)
    (rho c = 1), and the multiplicative perturbation 1 + 0.1 N(0,1) of w
    that generates the synthetic target."""
    rng = np.random.default_rng(seed)
    m = n + 2
    w = 0.5 * np.exp(0.3 * rng.standard_normal((m, m, m)))
    pert = 1.0 + 0.1 * rng.standard_normal((m, m, m))
    x = np.arange(-1, n + 1) / (n - 1.0)
    d2 = (x - 0.5) ** 2
    q = amp * np.exp(-(d2[:, None, None] + d2[None, :, None] + d2[None, None, :]) / (2 * width ** 2))
    h = 1.0 / (n - 1)
    tau = 0.9 / (6 * k / h ** 2 + float(w.max()))
    return {"w": w.ravel(), "q": q.ravel(), "pert": pert.ravel(), "tau": tau}


def heatvar_init_numpy(n, seed=0, amp=1e-3):
    """Node arrays (z,y,x) of the variable-kappa heat workload:
    u = prod sin(pi x_p) + amp*(2U-1) on inner nodes, kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y).
    Coordinates x_p = i/(n_p-1) on the unit cube."""
    nx, ny, nz = n
    i = np.arange(nx)
    j = np.arange(ny)
    k = np.arange(nz)
    x = (1.0 / (nx - 1)) * i
    y = (1.0 / (ny - 1)) * j
    z = (1.0 / (nz - 1)) * k
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    t = I.astype(np.uint64) + np.uint64(nx) * (J.astype(np.uint64) + np.uint64(ny) * K.astype(np.uint64))
    un = (splitmix64(seed, t) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    inner = (I > 0) & (I < nx - 1) & (J > 0) & (J < ny - 1) & (K > 0) & (K < nz - 1)
    u = np.where(inner, v + amp * (2.0 * un - 1.0), v)
    kap = 1.0 + 0.5 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    return u, kap


def sphere_sigma(n, center, radius, inside=3.546e7, h=1e-8):
    """Conductivity of the EMFDTD example's initInner (EMFDTD.hpp:258-276:
    the IndicatorEllipsoid of geometry/IndicatorEllipsoid.hpp:48-68 at the
    node indices) for equal radii, vectorised: (n, n, n) array indexed
    [k, j, i]."""
    ax = np.arange(n, dtype=np.float64)
    fz, fy, fx = np.meshgrid(ax - center[2], ax - center[1], ax - center[0], indexing="ij")
    ro = np.sqrt((fx * fx + fy * fy) + fz * fz)
    rn = float(radius)
    return 0.5 * inside * ((ro + rn) / np.sqrt((ro + rn) ** 2 + h * h) - (ro - rn) / np.sqrt((ro - rn) ** 2 + h * h))
