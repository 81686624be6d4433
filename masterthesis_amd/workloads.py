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
