"""torch.autograd.Function over the adjoint gradients of the C ABI.

PennesMisfit: J(w) = 0.5 * sum_inner (T^N(w) - target)^2 for the Pennes
bioheat forward model with perfusion field w; backward returns dJ/dw from
the reverse-time adjoint with binomial checkpointing (scafes_pennes_gradient).
The FP32 mode streams the deviation theta = T - T_a on the device.  With the
Tumor example's Neumann/Cauchy borders (TumorHeatEqnFDM.hpp:272-349) every
node is updated, and J and dJ/dw cover every node.

KappaMisfit: J(kappa) = 0.5 * sum_inner (u^N(kappa) - target)^2 for the
variable-kappa heat; backward returns dJ/dkappa (scafes_heat7var_gradient).

PermittivityMisfit: J(eps) = 0.5 * sum (E^N(eps) - target)^2 over the E
entries the Yee step updates; backward returns dJ/deps
(scafes_yee_eps_gradient).

Tensors are host (CPU) tensors: node fields in the memAll layout (ghost
width 1, shape (z, y, x) with n+2 per axis), Yee fields as (N+1)^3 arrays
per component; the device work runs on the given capi.Context.
"""
import numpy as np
import torch

from . import capi


class _PennesJ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, t0, q, target, runner):
        J, grad = runner._run(w.detach().cpu().numpy(), t0.detach().cpu().numpy(), q.detach().cpu().numpy(),
                              target.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(grad))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None, None


class PennesMisfit:
    """Callable J(w, T0, Q, target) with an adjoint backward pass."""

    def __init__(self, context, n, tau, cdiff, cperf, csrc, t_a, nsteps, snapshots=16, dtype="f32_kahan",
                 tumor=None, h=None):
        """dtype: "f64", "f32" or "f32_kahan" (FP32 fields with a compensated forward).
        tumor: None (border nodes keep their values) or (k_cond, h_conv, t_inf,
        q_bc) of the Tumor example's borders.  h: grid spacing per axis
        (default 1/(n_p - 1))."""
        self.ctx = context
        self.n = tuple(n)
        self.nsteps = int(nsteps)
        self.snapshots = int(snapshots)
        self.f64 = dtype == "f64"
        self.code = {"f64": capi.F64, "f32": capi.F32, "f32_kahan": capi.F32_KAHAN}[dtype]
        self.t_a = float(t_a)
        self.shift = 0.0 if self.f64 else self.t_a
        self.params = capi.PennesParams(tau=tau, cdiff=cdiff, cperf=cperf, csrc=csrc, t_a=self.t_a - self.shift)
        if tumor is not None:
            k_cond, h_conv, t_inf, q_bc = (float(v) for v in tumor)
            self.params.border_mode = capi.BORDER_TUMOR
            self.params.k_cond, self.params.h_conv, self.params.q_bc = k_cond, h_conv, q_bc
            self.params.t_inf = t_inf - self.shift
        self.geom = capi.geom(self.n, h=h)
        self.last_stats = None

    def _run(self, w, t0, q, target):
        dt = np.float64 if self.f64 else np.float32
        code = self.code
        bufs = [self.ctx.upload(np.ascontiguousarray(a, dtype=dt))
                for a in (t0 - self.shift, w, q, target - self.shift)]
        g = self.ctx.alloc(t0.size, np.float64)
        J = capi.C.c_double(0.0)
        stats = np.zeros(4)
        try:
            capi.check(self.ctx.L.scafes_pennes_gradient(
                self.ctx.h, capi.C.byref(self.geom), 1, code, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
                capi.C.byref(self.params), self.nsteps, self.snapshots, g.ptr, capi.C.byref(J),
                stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double))))
            grad = g.read().reshape(t0.shape)
        finally:
            for b in bufs + [g]:
                b.free()
        self.last_stats = stats
        return J.value, grad

    def __call__(self, w, t0, q, target):
        return _PennesJ.apply(w, t0, q, target, self)


class _KappaJ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, u0, target, runner):
        J, grad = runner._run(kappa.detach().cpu().numpy(), u0.detach().cpu().numpy(),
                              target.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(grad))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None


class KappaMisfit:
    """Callable J(kappa, u0, target) of the variable-kappa heat with an
    adjoint backward pass."""

    def __init__(self, context, n, tau, nsteps, snapshots=16, dtype="f32_kahan", border="keep",
                 border_value=0.0):
        """dtype: "f64", "f32" or "f32_kahan"; border: "keep" or "value"."""
        self.ctx = context
        self.n = tuple(n)
        self.tau = float(tau)
        self.nsteps = int(nsteps)
        self.snapshots = int(snapshots)
        self.mode = {"f64": capi.HEATVAR_F64, "f32": capi.HEATVAR_F32, "f32_kahan": capi.HEATVAR_F32_KAHAN}[dtype]
        self.border = {"keep": capi.BORDER_KEEP, "value": capi.BORDER_VALUE}[border]
        self.border_value = float(border_value)
        self.geom = capi.geom(self.n)
        self.last_stats = None

    def _run(self, kappa, u0, target):
        dt = np.float64 if self.mode == capi.HEATVAR_F64 else np.float32
        bufs = [self.ctx.upload(np.ascontiguousarray(a, dtype=dt)) for a in (u0, kappa, target)]
        g = self.ctx.alloc(u0.size, np.float64)
        J = capi.C.c_double(0.0)
        stats = np.zeros(4)
        try:
            capi.check(self.ctx.L.scafes_heat7var_gradient(
                self.ctx.h, capi.C.byref(self.geom), 1, self.mode, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, self.tau,
                self.border, self.border_value, self.nsteps, self.snapshots, g.ptr, capi.C.byref(J),
                stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double))))
            grad = g.read().reshape(u0.shape)
        finally:
            for b in bufs + [g]:
                b.free()
        self.last_stats = stats
        return J.value, grad

    def __call__(self, kappa, u0, target):
        return _KappaJ.apply(kappa, u0, target, self)


class _PermittivityJ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, f0, target, runner):
        J, grad = runner._run(eps.detach().cpu().numpy(), f0.detach().cpu().numpy(),
                              target.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(grad))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None


class PermittivityMisfit:
    """Callable J(eps, f0, target) of the Yee leapfrog in a PEC box of ncell
    cells with a relative permittivity field eps ((N+1)^3), J = 0.5 sum over
    the updated E entries of (E^N - target)^2, with an adjoint backward
    pass.  f0: the six Yee arrays; target: the three E arrays."""

    def __init__(self, context, ncell, s, nsteps, snapshots=16, dtype="f32"):
        """dtype: "f64" or "f32" (FP32 fields, FP64 gradient)."""
        self.ctx = context
        self.ncell = int(ncell)
        self.s = float(s)
        self.nsteps = int(nsteps)
        self.snapshots = int(snapshots)
        self.code = {"f64": capi.F64, "f32": capi.F32}[dtype]
        self.last_stats = None

    def _run(self, eps, f0, target):
        dt = np.float64 if self.code == capi.F64 else np.float32
        bufs = [self.ctx.upload(np.ascontiguousarray(a, dtype=dt)) for a in (f0, eps, target)]
        g = self.ctx.alloc(eps.size, np.float64)
        J = capi.C.c_double(0.0)
        stats = np.zeros(4)
        try:
            capi.check(self.ctx.L.scafes_yee_eps_gradient(
                self.ctx.h, self.ncell, self.code, self.s, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, self.nsteps,
                self.snapshots, g.ptr, capi.C.byref(J), stats.ctypes.data_as(capi.C.POINTER(capi.C.c_double))))
            grad = g.read().reshape(eps.shape)
        finally:
            for b in bufs + [g]:
                b.free()
        self.last_stats = stats
        return J.value, grad

    def __call__(self, eps, f0, target):
        return _PermittivityJ.apply(eps, f0, target, self)
