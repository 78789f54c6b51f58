"""Differentiable implicit time stepping of the (H, Theta) solidification
problem: the hand-written adjoint of the backward-Euler step
(ml_hc_adjoint_step, GPU kernels + BiCGStab on the transposed step
Jacobian) behind a torch.autograd.Function, with the forward states
recomputed from checkpoints every `checkpoint` steps.

Parameters: the bottom heat-flux profile F (one value per bottom face), the
Stefan number St (closure), and a permeability prefactor.  Without flow the
(H, Theta) step does not depend on the permeability (its face coefficients
use the porosity only, srcSubcycle/AMRLevelMushyLayerInit.cpp:925-951), so
that gradient is exactly zero.

The forward step is the library's FP64 backward-Euler solve; its per-step
sequence is porosity -> bCoef -> set_coefs(1, -dt) -> heat-flux source ->
solve_be (multiCompAdvectDiffuse with the coefficients at t^n,
srcSubcycle/AMRLevelMushyLayer.cpp:1645-1831).
"""
import numpy as np
import torch

from . import (DeviceArray, HCOp, default_opts, face_shape, make_bc, make_params, valid)


class HeatFluxAdjoint:
    """One heat-flux workload (a `workloads.heat_flux` case) on a context."""

    def __init__(self, ctx, case, checkpoint=10, opts=None, adj_opts=None):
        self.ctx = ctx
        self.case = case
        self.n = list(case["n"])
        if len(self.n) != 2:
            raise ValueError("the adjoint workload is 2D")
        self.steps = case["steps"]
        self.dt = case["dt"]
        self.checkpoint = checkpoint
        # forward solves: to 1e-10 without the hang rule (the FAS residual
        # need not fall monotonically), then on from there until it stops
        # falling (the rounding floor): the adjoint assumes R(phi^n) = 0
        self.opts = opts or default_opts(max_iter=400, hang=-1.0)
        self.polish = default_opts(max_iter=200, eps=1e-15, norm_thresh=1e-15, hang=1e-10)
        self.adj_opts = adj_opts or default_opts(eps=1e-13, max_iter=1000)
        shp = case["phi"].shape
        self.shape = shp
        self.phi0 = DeviceArray(ctx, shp, host=case["phi"])
        self.chi = DeviceArray(ctx, (1,) + shp[1:])
        self.b = [DeviceArray(ctx, face_shape(self.n, d, 2)) for d in range(2)]
        self.old = DeviceArray(ctx, shp)
        self.src = DeviceArray(ctx, shp)
        self.flux = DeviceArray(ctx, (self.n[0],))
        self._ops = {}
        self.adj_stats = []

    def op(self, St):
        if St not in self._ops:
            p = dict(self.case["params"])
            p["stefan"] = St
            self._ops[St] = HCOp(self.ctx, self.n, self.case["dx"], self.case["periodic"], make_params(**p),
                                 make_bc(self.case["bc"], apply=1))
        return self._ops[St]

    def advance(self, op, phi):
        """One implicit step of phi in place (self.flux holds F)."""
        op.porosity(phi, self.chi)
        op.assemble_bcoef(self.chi, self.b)
        op.set_coefs(None, self.b, 1.0, -self.dt)
        op.flux_source(self.flux, self.src)
        self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, self.old.ptr, phi.ptr, phi.nbytes), "copy")
        st = op.solve_be(phi, self.old, self.dt, src=self.src, opts=self.opts)
        if not st.exit_status & (1 | 8):
            raise RuntimeError("forward step did not converge (status %d)" % st.exit_status)
        op.solve_be(phi, self.old, self.dt, src=self.src, opts=self.polish)
        return st

    def _copy(self, x):
        y = DeviceArray(self.ctx, x.shape)
        self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, y.ptr, x.ptr, x.nbytes), "copy")
        return y

    def forward(self, F, St):
        """Final state (host array) and its porosity; device checkpoints of
        the states at steps 0, k, 2k, ..."""
        op = self.op(St)
        self.flux.set(np.asarray(F, dtype=np.float64))
        phi = self._copy(self.phi0)
        cps = {0: self._copy(phi)}
        for n in range(1, self.steps + 1):
            self.advance(op, phi)
            if n % self.checkpoint == 0 and n < self.steps:
                cps[n] = self._copy(phi)
        op.porosity(phi, self.chi)
        out = phi.get(), self.chi.get()
        phi.free()
        return out[0], out[1], cps

    def backward(self, F, St, cps, g_phi, g_chi):
        """dL/dF and dL/dSt from dL/dphi_N and dL/dchi_N (host arrays)."""
        op = self.op(St)
        self.flux.set(np.asarray(F, dtype=np.float64))
        gF = DeviceArray(self.ctx, (self.n[0],), host=np.zeros(self.n[0]))
        g = DeviceArray(self.ctx, self.shape, host=np.zeros(self.shape))
        gold = DeviceArray(self.ctx, self.shape, host=np.zeros(self.shape))
        gst = 0.0
        self.adj_stats = []
        starts = sorted(cps)
        for s0 in reversed(starts):
            s1 = min(s0 + self.checkpoint, self.steps)
            states = [self._copy(cps[s0])]
            for _ in range(s0 + 1, s1 + 1):
                x = self._copy(states[-1])
                self.advance(op, x)
                states.append(x)
            if s1 == self.steps:
                gp = np.zeros(self.shape)
                valid(self.n, gp)[:] = valid(self.n, np.asarray(g_phi, dtype=np.float64))
                g.set(gp)
                gc = np.zeros((1,) + self.shape[1:])
                valid(self.n, gc)[:] = valid(self.n, np.asarray(g_chi, dtype=np.float64))
                gcd = DeviceArray(self.ctx, gc.shape, host=gc)
                gst += op.porosity_vjp(states[-1], gcd, g)
                gcd.free()
            for k in range(len(states) - 1, 0, -1):
                dst, st = op.adjoint_step(states[k], states[k - 1], self.dt, g, gold, gF, opts=self.adj_opts)
                if st.exit_status != 1 and st.exit_status != 8:
                    raise RuntimeError("adjoint solve did not converge (status %d, %d iterations)"
                                       % (st.exit_status, st.iterations))
                self.adj_stats.append((st.iterations, st.final_residual / max(st.initial_residual, 1e-300)))
                gst += dst
                g, gold = gold, g
            for x in states:
                x.free()
        out = gF.get()
        for a in (gF, g, gold):
            a.free()
        return out, gst

    def close(self):
        for op in self._ops.values():
            op.close()
        self._ops = {}


class ImplicitSteps(torch.autograd.Function):
    """(flux[nx], permeability prefactor, St) -> (final state, its porosity),
    both FAB-shaped float64 tensors on the flux's device."""

    @staticmethod
    def forward(ctx, flux, perm, stefan, model):
        F = flux.detach().to("cpu", torch.float64).numpy()
        St = float(stefan)
        phi, chi, cps = model.forward(F, St)
        ctx.model, ctx.F, ctx.St, ctx.cps = model, F, St, cps
        ctx.perm_shape = perm.shape
        dev = flux.device
        return torch.from_numpy(phi).to(dev), torch.from_numpy(chi).to(dev)

    @staticmethod
    def backward(ctx, g_phi, g_chi):
        gp = g_phi.detach().to("cpu", torch.float64).numpy() if g_phi is not None else np.zeros(ctx.model.shape)
        gc = (g_chi.detach().to("cpu", torch.float64).numpy() if g_chi is not None
              else np.zeros((1,) + ctx.model.shape[1:]))
        gF, gSt = ctx.model.backward(ctx.F, ctx.St, ctx.cps, gp, gc)
        for x in ctx.cps.values():
            x.free()
        dev = g_phi.device if g_phi is not None else g_chi.device
        return (torch.from_numpy(gF).to(dev), torch.zeros(ctx.perm_shape, dtype=torch.float64, device=dev),
                torch.tensor(gSt, dtype=torch.float64, device=dev), None)


def field_loss(n, phi, chi, target_phi, target_chi):
    """Mean squared misfit of the Theta and porosity fields (valid cells)."""
    t = torch.as_tensor
    th = phi[1, :, 1:-1, 1:-1]
    x = chi[0, :, 1:-1, 1:-1]
    return (torch.mean((th - t(target_phi, device=phi.device)[1, :, 1:-1, 1:-1]) ** 2) +
            torch.mean((x - t(target_chi, device=phi.device)[0, :, 1:-1, 1:-1]) ** 2))
