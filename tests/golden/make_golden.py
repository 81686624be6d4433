"""Writes the golden fixtures of this directory: small seeded inputs and the
FP64 outputs of every kernel family on the path, one .npz per case (plain
arrays, loadable with numpy.load(allow_pickle=False)), and manifest.json with
each case's parameters and the SHA-256 of its file.

The expected outputs come from the oracle (oracle/, the C restatement of the
reference's CPU path), after it has been pinned by the reference's
known-answer problems and by closed forms and invariants
(tests/test_oracle_*.py).  The reference's own tree holds no numeric fixtures
for this path, and it cannot be built here without stand-ins, so these files
freeze the pinned oracle: tests/test_golden_cpu.py checks that the oracle
still reproduces them bit for bit, and tests/test_gpu_golden.py checks the
HIP kernels against the files directly.

Run from the repository root (rewrites every file):
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def sine_mode(g, n):
    x = np.linspace(0.0, 1.0, n)
    s = np.sin(np.pi * x)
    u = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u)[...] = s[:, None, None] * s[None, :, None] * s[None, None, :]
    return u


def case_heat7_sine():
    """Constant-coefficient heat, 16^3, Dirichlet 0, sine mode, 20 steps."""
    n, N = 16, 20
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    tau = 0.9 * h * h / 6.0
    u0 = sine_mode(g, n)
    u = u0.copy()
    coef = O.heat7_coef(tau)
    for _ in range(N):
        u = O.heat7_step(g, u, border=O.BORDER_VALUE, border_value=0.0, coef=coef)
    return {"n": n, "nsteps": N, "tau": tau}, {"u0": u0, "u": u}


def case_heat7var():
    """Variable-kappa conservative heat, 16^3, kappa = 1 + 0.5 sin(2 pi x)
    sin(2 pi y), sine mode + seeded noise, border kept, 12 steps."""
    n, N = 16, 12
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(3)
    u0 = sine_mode(g, n)
    O.nodes_view(g, u0)[...] += 1e-3 * rng.uniform(-1.0, 1.0, (n, n, n))
    x = np.linspace(0.0, 1.0, n)
    kap = np.ones(O.memall_shape(g))
    O.nodes_view(g, kap)[...] = 1.0 + 0.5 * np.sin(2 * np.pi * x)[None, None, :] * np.sin(2 * np.pi * x)[None, :, None]
    tau = 0.9 * h * h / (6.0 * 1.5)
    u = O.heat7var_run(g, u0, kap, tau, N, border=O.BORDER_KEEP)
    return {"n": n, "nsteps": N, "tau": tau}, {"u0": u0, "kappa": kap, "u": u}


def case_pennes():
    """Pennes with perfusion and source fields, 12^3, 10 steps: final state
    and the gradient of J = 0.5 sum_inner (T^N - target)^2 w.r.t. w."""
    n, N = 12, 10
    g = O.make_grid((n, n, n))
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(11)
    w = 0.5 * np.exp(0.3 * rng.standard_normal(O.memall_shape(g)))
    x = np.linspace(0, 1, n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    q = np.zeros(O.memall_shape(g))
    O.nodes_view(g, q)[...] = 50.0 * np.exp(-((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) / (2 * .15 ** 2))
    tau = 0.9 / (6 * 0.5 / h ** 2 + w.max())
    coef = O.pennes_coef(tau=tau, cdiff=0.5, cperf=1.0, csrc=1.0, t_a=37.0)
    T0 = np.full(O.memall_shape(g), 37.0)
    w_pert = w * (1 + 0.2 * rng.standard_normal(w.shape))
    target = O.pennes_forward_all(g, T0, w_pert, q, coef, N)[-1]
    st = O.pennes_forward_all(g, T0, w, q, coef, N)
    J, grad = O.pennes_grad(g, st, w, target, coef, N)
    return ({"n": n, "nsteps": N, "tau": tau, "cdiff": 0.5, "cperf": 1.0, "csrc": 1.0, "t_a": 37.0, "J": J},
            {"T0": T0, "w": w, "q": q, "target": target, "T": st[-1], "grad": grad})


def case_yee():
    """Yee leapfrog, PEC box of 12^3 cells, Courant 0.5, random valid
    fields, 10 steps."""
    N, steps, s = 12, 10, 0.5
    rng = np.random.default_rng(12)
    f = O.yee_alloc(N)
    f[0, 1:N, 1:N, :N] = rng.standard_normal((N - 1, N - 1, N))
    f[1, 1:N, :N, 1:N] = rng.standard_normal((N - 1, N, N - 1))
    f[2, :N, 1:N, 1:N] = rng.standard_normal((N, N - 1, N - 1))
    f[3, :N, :N, :] = 0.1 * rng.standard_normal((N, N, N + 1))
    f[4, :N, :, :N] = 0.1 * rng.standard_normal((N, N + 1, N))
    f[5, :, :N, :N] = 0.1 * rng.standard_normal((N + 1, N, N))
    f0 = f.copy()
    O.yee_step(N, s, f, steps)
    return {"ncell": N, "nsteps": steps, "courant": s}, {"f0": f0, "f": f}


def case_yee_eps():
    """Yee with a relative permittivity field, PEC box of 10^3 cells, Courant
    0.5, random valid fields, eps in [1, 4), 8 steps: final state and the
    gradient of J = 0.5 sum over the updated E entries of (E^N - target)^2
    w.r.t. eps."""
    N, steps, s = 10, 8, 0.5
    rng = np.random.default_rng(13)
    m = N + 1
    f0 = rng.standard_normal((6, m, m, m))
    f0[:3] *= O.yee_emask(N)
    eps = 1.0 + 3.0 * rng.random((m, m, m))
    target = 0.3 * rng.standard_normal((3, m, m, m))
    st = O.yee_eps_forward_all(N, s, eps, f0, steps)
    J, grad = O.yee_eps_grad(N, s, eps, st, target, steps)
    return ({"ncell": N, "nsteps": steps, "courant": s, "J": J},
            {"f0": f0, "eps": eps, "target": target, "f": st[-1], "grad": grad})


def case_heat2d():
    """2-D heat problems of the batched ensemble (n = 9, 17, 33), 40 steps:
    final fields, J and dJ/dkappa."""
    sizes, N = [9, 17, 33], 40
    rng = np.random.default_rng(5)
    out = {}
    kap, taus, Js, grads = [], [], [], []
    for n in sizes:
        # kappa and tau exactly representable in FP32 (the batched kernel's precision)
        k = float(np.float32(rng.uniform(0.1, 1.0)))
        h = 1.0 / (n - 1)
        tau = float(np.float32(0.9 * h * h / 4.0))
        x = np.linspace(0.0, 1.0, n)
        u0 = np.outer(np.sin(np.pi * x), np.sin(np.pi * x))
        u0[1:-1, 1:-1] += 0.01 * rng.standard_normal((n - 2, n - 2))
        u0[0, :] = u0[-1, :] = u0[:, 0] = u0[:, -1] = 0.0
        tgt = u0 * np.exp(-2 * np.pi ** 2 * 1.1 * k * N * tau)
        u, _ = O.heat2d_run(n, k, tau, N, u0)
        J, d = O.heat2d_grad(n, k, tau, N, u0, tgt)
        out[f"u0_{n}"], out[f"target_{n}"], out[f"u_{n}"] = u0, tgt, u
        kap.append(k)
        taus.append(tau)
        Js.append(J)
        grads.append(d)
    out["kappa"], out["tau"], out["J"], out["dJ_dkappa"] = (np.array(v) for v in (kap, taus, Js, grads))
    return {"sizes": sizes, "nsteps": N}, out


def case_tumor():
    """Tumor example borders (Neumann faces, Cauchy far face), 13x9x8 on
    0.06 x 0.03 x 0.03 m, random initial state, 12 steps."""
    n, N = (13, 9, 8), 12
    ext = (0.06, 0.03, 0.03)
    g = O.make_grid(n, x0=[0.0] * 3, x1=list(ext))
    h = [ext[p] / (n[p] - 1) for p in range(3)]
    rho, cc, K, H = 1040.0, 3650.0, 0.6, 10.0
    tau = 0.5 / (sum(2.0 * K / (rho * cc) / hp ** 2 for hp in h) + 2.0 * H / (rho * cc * h[-1]) +
                 1052.0 * 3800.0 / (rho * cc) * 0.004)
    coef = O.tumor_coef(tau, [0.03, 0.015, 0.022], 0.006, q_bc=120.0)
    w, q = O.tumor_fields(g, coef)
    rng = np.random.default_rng(13)
    u0 = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u0)[...] = 37.0 + rng.standard_normal(tuple(reversed(n)))
    u = u0
    for _ in range(N):
        u = O.tumor_step(g, u, coef)
    par = {k: getattr(coef, k) for k, _ in coef._fields_ if k != "center"}
    par.update({"n": list(n), "extent": list(ext), "nsteps": N, "center": list(coef.center)})
    return par, {"u0": u0, "w": w, "q": q, "u": u}


def case_mri():
    """MRI-data example: nine random material fields with a ragged open-skull
    patch, 11x9x8, radiation on, 12 steps."""
    n, N = (11, 9, 8), 12
    ext = (0.06, 0.03, 0.03)
    g = O.make_grid(n, x0=[0.0] * 3, x1=list(ext))
    h = [ext[p] / (n[p] - 1) for p in range(3)]
    rng = np.random.default_rng(14)
    shape = tuple(reversed(n))
    mats = [1040.0 * (1 + 0.1 * rng.random(shape)), 3650.0 * (1 + 0.1 * rng.random(shape)),
            0.5 + 0.2 * rng.random(shape), np.full(shape, 1052.0), np.full(shape, 3800.0),
            0.004 + 0.006 * rng.random(shape), 35.0 + rng.random(shape), 25000.0 * rng.random(shape),
            np.zeros(shape)]
    mats[8][-1] = rng.random(mats[8][-1].shape) < 0.4
    tau = 0.4 / sum(2.0 * 0.7 / (1040.0 * 3650.0) / hp ** 2 for hp in h)
    mc = O.mri_coef(tau, h=10.0, t_amb=22.0, q_bc=35.0, q_skull=-12.0, epsilon=0.98)
    u0 = np.zeros(O.memall_shape(g))
    O.nodes_view(g, u0)[...] = 36.0 + rng.standard_normal(shape)
    u = u0
    for _ in range(N):
        u = O.mri_step(g, u, mats, mc)
    par = {"n": list(n), "extent": list(ext), "nsteps": N, "tau": tau, "h_conv": 10.0, "t_amb": 22.0,
           "q_bc": 35.0, "q_skull": -12.0, "epsilon": 0.98, "sigma": 5.670373e-8}
    arrs = {"u0": u0, "u": u}
    for name, m in zip(O.MRI_FIELDS, mats):
        arrs["mat_" + name] = m
    return par, arrs


def case_emfdtd():
    """EMFDTD example: 16^3 nodes, CPML depth 3, conductive sphere, hard
    source and sink, 30 leapfrog steps: the 24 state fields and the sink
    series."""
    n, d, N = 16, 3, 30
    src, sink = (7, 7, 7), (7, 7, 10)
    sigma = O.ellipsoid_sigma(n, (11, 8, 8), (2.5,) * 3)
    ref = O.Emfdtd(n, d=d, src=src, sink=sink, sigma=sigma)
    state0 = ref.f.copy()
    s = ref.run(N)
    par = {"n": n, "d": d, "nsteps": N, "src": list(src), "sink": list(sink), "cfln": 0.5}
    return par, {"state0": state0, "state": ref.f[:O.EM_STATE].copy(), "sink": s}


CASES = {
    "heat7_sine_16": case_heat7_sine,
    "heat7var_16": case_heat7var,
    "pennes_grad_12": case_pennes,
    "yee_12": case_yee,
    "yee_eps_10": case_yee_eps,
    "heat2d_batch": case_heat2d,
    "tumor_13x9x8": case_tumor,
    "mri_11x9x8": case_mri,
    "emfdtd_16": case_emfdtd,
}


def sha256(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    manifest = {}
    for name, fn in CASES.items():
        par, arrs = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **{k: np.ascontiguousarray(v, dtype=np.float64) for k, v in arrs.items()})
        manifest[name] = {"params": par, "arrays": sorted(arrs), "sha256": sha256(path)}
        print(f"{name}: {os.path.getsize(path)} bytes")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
