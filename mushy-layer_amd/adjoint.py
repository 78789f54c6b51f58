"""Differentiable implicit time stepping of the (H, Theta) solidification
problem: the hand-written adjoint of the backward-Euler step
(ml_hc_adjoint_step, GPU kernels + BiCGStab on the transposed step
Jacobian, preconditioned by nu1 sweeps of the transposed smoother -- one by
default) behind a torch.autograd.Function, with the forward states
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
import weakref

import numpy as np
import torch

from . import (DeviceArray, HCOp, default_opts, face_shape, make_bc, make_params, valid)


def _free_arrays(arrays):
    for a in arrays:
        try:
            a.free()
        except Exception:   # the context may already be closed at interpreter exit
            pass


class Checkpoints:
    """Device checkpoints of one forward pass (states at steps 0, k, 2k, ...).
    They live as long as this object -- the autograd graph that holds it --
    so backward may run more than once (retain_graph=True); the memory is
    freed when the graph is released, or by close()."""

    def __init__(self, states):
        self.states = states
        self._fin = weakref.finalize(self, _free_arrays, list(states.values()))

    def close(self):
        self._fin()
        self.states = {}

    def check(self):
        if not self.states:
            raise RuntimeError("the forward checkpoints were released (Checkpoints.close)")


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

    def set_flux(self, F):
        """F: host values, or (device address, nbytes) of a float64 device buffer."""
        if isinstance(F, tuple):
            self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, self.flux.ptr, F[0], F[1]), "copy")
        else:
            self.flux.set(np.asarray(F, dtype=np.float64))

    def forward_device(self, F, St):
        """Final state and its porosity as device arrays, and the device
        checkpoints (a Checkpoints holder)."""
        op = self.op(St)
        self.set_flux(F)
        phi = self._copy(self.phi0)
        cps = {0: self._copy(phi)}
        for n in range(1, self.steps + 1):
            self.advance(op, phi)
            if n % self.checkpoint == 0 and n < self.steps:
                cps[n] = self._copy(phi)
        chi = DeviceArray(self.ctx, (1,) + self.shape[1:])
        op.porosity(phi, chi)
        return phi, chi, Checkpoints(cps)

    def forward(self, F, St):
        """Final state (host array) and its porosity; device checkpoints of
        the states at steps 0, k, 2k, ..."""
        phi, chi, cps = self.forward_device(F, St)
        out = phi.get(), chi.get()
        phi.free()
        chi.free()
        return out[0], out[1], cps

    def backward(self, F, St, cps, g_phi, g_chi, device_out=False):
        """dL/dF and dL/dSt from dL/dphi_N and dL/dchi_N: host arrays, or
        device addresses of FAB-shaped float64 buffers whose ghost cells are
        zero.  dL/dF is a host array, or a DeviceArray with device_out."""
        op = self.op(St)
        if isinstance(cps, Checkpoints):
            cps.check()
            cps = cps.states
        self.set_flux(F)
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
                gcd = DeviceArray(self.ctx, (1,) + self.shape[1:])
                if isinstance(g_phi, int):
                    self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, g.ptr, g_phi, g.nbytes), "copy")
                    self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, gcd.ptr, g_chi, gcd.nbytes), "copy")
                else:
                    gp = np.zeros(self.shape)
                    valid(self.n, gp)[:] = valid(self.n, np.asarray(g_phi, dtype=np.float64))
                    g.set(gp)
                    gc = np.zeros((1,) + self.shape[1:])
                    valid(self.n, gc)[:] = valid(self.n, np.asarray(g_chi, dtype=np.float64))
                    gcd.set(gc)
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
        for a in (g, gold):
            a.free()
        if device_out:
            return gF, gst
        out = gF.get()
        gF.free()
        return out, gst

    def close(self):
        for op in self._ops.values():
            op.close()
        self._ops = {}


class ImplicitSteps(torch.autograd.Function):
    """(flux[nx], permeability prefactor, St) -> (final state, its porosity),
    both FAB-shaped float64 tensors on the flux's device.

    The states, checkpoints and adjoint iterates stay in the library's device
    memory; only the flux profile, the final state, its porosity and the
    gradients cross this boundary, as host arrays.  (torch's wheel carries
    its own HIP/HSA runtime, distinct from the system ROCm runtime the
    library links, and two HSA runtimes cannot both open the GPU in one
    process, so torch device tensors are not exchanged.)  The gradient with
    respect to the permeability prefactor is structurally zero: without flow
    the (H, Theta) step does not depend on the permeability (the HC face
    coefficients use the porosity only), so the returned zeros are exact,
    not an approximation.  The forward checkpoints live with the graph
    (Checkpoints), so backward may be repeated (retain_graph=True)."""

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
        model = ctx.model
        gp = g_phi.detach().to("cpu", torch.float64).numpy() if g_phi is not None else np.zeros(model.shape)
        gc = (g_chi.detach().to("cpu", torch.float64).numpy() if g_chi is not None
              else np.zeros((1,) + model.shape[1:]))
        gF, gSt = model.backward(ctx.F, ctx.St, ctx.cps, gp, gc)
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
