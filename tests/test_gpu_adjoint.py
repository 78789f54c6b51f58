"""Adjoint of the implicit (H, Theta) steps: the hand-written GPU adjoint
(ml_hc_adjoint_step behind a torch.autograd.Function, states recomputed from
checkpoints) against central finite differences of the same GPU forward model
and of the FP64 oracle's forward model, on a 32x32 heat-flux case (12 steps,
checkpoints every 5, mush forming at the cooled bottom).

Bars: the forward state equals the oracle's to 1e-12 relative; directional
derivatives agree with central differences (h = 1e-6 relative to the
parameter's scale) to FD_TOL relative; the permeability prefactor's gradient
is exactly zero (no flow: the step does not use the permeability)."""
import importlib

import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

FD_TOL = 2e-5
H_REL = 1e-6


def _setup(ml, gpu_ctx, wl):
    torch = pytest.importorskip("torch")
    adj = importlib.import_module("mushy-layer_amd.adjoint")
    case = wl.heat_flux(n=(32, 32), steps=12, seed=9)
    model = adj.HeatFluxAdjoint(gpu_ctx, case, checkpoint=5)
    F_true = wl.flux_profile(32, F0=12.0, modes=((1, 0.3, 0.0), (2, 0.1, 1.0)))
    tphi, tchi, cps = model.forward(F_true, 5.5)
    cps.close()
    return torch, adj, case, model, tphi, tchi


def _loss_np(n, phi, chi, tphi, tchi):
    v = lambda a: a[..., 1:-1, 1:-1]
    return float(np.mean((v(phi)[1] - v(tphi)[1]) ** 2) + np.mean((v(chi)[0] - v(tchi)[0]) ** 2))


def _gpu_loss(model, F, St, tphi, tchi):
    phi, chi, cps = model.forward(F, St)
    cps.close()
    return _loss_np(model.n, phi, chi, tphi, tchi)


def _adjoint_grads(torch, adj, model, F, St, tphi, tchi):
    flux = torch.tensor(F, dtype=torch.float64, requires_grad=True)
    perm = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    st = torch.tensor(St, dtype=torch.float64, requires_grad=True)
    phi, chi = adj.ImplicitSteps.apply(flux, perm, st, model)
    loss = adj.field_loss(model.n, phi, chi, tphi, tchi)
    loss.backward()
    return float(loss.detach()), flux.grad.numpy(), float(perm.grad), float(st.grad)


def test_adjoint_matches_gpu_finite_differences(ml, gpu_ctx, wl):
    torch, adj, case, model, tphi, tchi = _setup(ml, gpu_ctx, wl)
    F = wl.flux_profile(32, F0=10.0)
    St = 5.0
    L, gF, gperm, gSt = _adjoint_grads(torch, adj, model, F, St, tphi, tchi)
    print("loss %.6e, |dL/dF| %.3e, dL/dSt %.6e, adjoint solves (iterations, rel residual) %s"
          % (L, np.linalg.norm(gF), gSt, model.adj_stats[:3]))
    assert L > 0
    assert gperm == 0.0
    assert all(r <= 1e-12 for _, r in model.adj_stats)
    rng = np.random.default_rng(3)
    for k in range(2):
        v = rng.standard_normal(F.shape)
        v /= np.linalg.norm(v)
        h = H_REL * np.max(np.abs(F))
        fd = (_gpu_loss(model, F + h * v, St, tphi, tchi) - _gpu_loss(model, F - h * v, St, tphi, tchi)) / (2 * h)
        ad = float(gF @ v)
        print("direction %d: adjoint %.9e, central difference %.9e" % (k, ad, fd))
        assert abs(ad - fd) <= FD_TOL * abs(fd), (k, ad, fd)
    h = H_REL * St
    fd = (_gpu_loss(model, F, St + h, tphi, tchi) - _gpu_loss(model, F, St - h, tphi, tchi)) / (2 * h)
    print("St: adjoint %.9e, central difference %.9e" % (gSt, fd))
    assert abs(gSt - fd) <= FD_TOL * abs(fd), (gSt, fd)
    model.close()


def _oracle_forward(case, F, St):
    n = case["n"]
    c = dict(case)
    c["params"] = dict(case["params"], stefan=St)
    c["apply_bc"] = 1
    g, p, bc = HC.oracle_setup(c)
    phi = case["phi"].copy()
    # the forward step of the adjoint model: to 1e-10 without the hang rule,
    # then on to the rounding floor
    opts = O.default_opts(max_iter=400, hang=-1.0)
    polish = O.default_opts(max_iter=200, eps=1e-15, norm_thresh=1e-15, hang=1e-10)
    for _ in range(case["steps"]):
        chi = O.porosity_field(g, p, phi)
        b = O.hc_bcoef(g, p, chi)
        src = np.zeros_like(phi)
        src[0, 0, 1, 1:-1] = b[1][0, 0, 0, :] * F / case["dx"]
        old = phi.copy()
        st = O.hc_be_step(g, p, bc, None, b, case["dt"], phi, old, src=src, opts=opts)
        assert st.exit_status & (1 | 8)
        O.hc_be_step(g, p, bc, None, b, case["dt"], phi, old, src=src, opts=polish)
    return phi, O.porosity_field(g, p, phi)


def test_adjoint_matches_oracle_finite_differences(ml, gpu_ctx, wl):
    """The reference FP64 gradient: central differences of the oracle's
    forward model (C, FP64) against the GPU adjoint."""
    torch, adj, case, model, tphi, tchi = _setup(ml, gpu_ctx, wl)
    F = wl.flux_profile(32, F0=10.0)
    St = 5.0
    phi_gpu, _, cps = model.forward(F, St)
    cps.close()
    phi_ref, _ = _oracle_forward(case, F, St)
    n = case["n"]
    err = HC.rel_err(HC.valid(n, phi_gpu), HC.valid(n, phi_ref))
    print("forward state, GPU vs oracle: %.3e" % err)
    assert err <= 1e-12
    L, gF, _, gSt = _adjoint_grads(torch, adj, model, F, St, tphi, tchi)

    def loss(FF, SS):
        ph, ch = _oracle_forward(case, FF, SS)
        return _loss_np(n, ph, ch, tphi, tchi)

    v = np.cos(2 * np.pi * (np.arange(32) + 0.5) / 32)
    v /= np.linalg.norm(v)
    h = H_REL * np.max(np.abs(F))
    fd = (loss(F + h * v, St) - loss(F - h * v, St)) / (2 * h)
    print("oracle direction: adjoint %.9e, central difference %.9e" % (float(gF @ v), fd))
    assert abs(float(gF @ v) - fd) <= FD_TOL * abs(fd)
    h = H_REL * St
    fd = (loss(F, St + h) - loss(F, St - h)) / (2 * h)
    print("oracle St: adjoint %.9e, central difference %.9e" % (gSt, fd))
    assert abs(gSt - fd) <= FD_TOL * abs(fd)
    model.close()


def test_adjoint_repeated_backward(ml, gpu_ctx, wl):
    """backward can be repeated (retain_graph): the checkpoints live with the
    graph, not with one backward; the permeability gradient is the
    structural zero."""
    torch, adj, case, model, tphi, tchi = _setup(ml, gpu_ctx, wl)
    F = wl.flux_profile(32, F0=10.0)
    flux = torch.tensor(F, dtype=torch.float64, requires_grad=True)
    perm = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    st = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    phi, chi = adj.ImplicitSteps.apply(flux, perm, st, model)
    loss = adj.field_loss(model.n, phi, chi, tphi, tchi)
    loss.backward(retain_graph=True)
    g1 = flux.grad.detach().numpy().copy()
    s1 = float(st.grad)
    assert float(perm.grad) == 0.0
    flux.grad = None
    st.grad = None
    loss.backward()
    np.testing.assert_array_equal(flux.grad.detach().numpy(), g1)
    assert float(st.grad) == s1
    assert np.abs(g1).max() > 0


def test_transposed_smoother_preconditioner(ml, gpu_ctx, wl):
    """The adjoint solves preconditioned by the transposed smoother (block
    red/black Gauss-Seidel on J^T, black first; one or two sweeps) give the
    gradients of the block-diagonal preconditioner to 1e-10 relative, in
    fewer iterations."""
    torch, adj, case, model, tphi, tchi = _setup(ml, gpu_ctx, wl)
    F = wl.flux_profile(32, F0=10.0)
    res = {}
    for sweeps in (0, 1, 2):
        model.adj_opts = ml.default_opts(eps=1e-13, max_iter=1000, nu1=sweeps)
        L, gF, gperm, gSt = _adjoint_grads(torch, adj, model, F, 5.0, tphi, tchi)
        its = [a for a, _ in model.adj_stats]
        assert all(r <= 1e-12 for _, r in model.adj_stats)
        res[sweeps] = (gF.copy(), gSt, float(np.mean(its)))
        print("sweeps %d: mean adjoint iterations %.2f" % (sweeps, res[sweeps][2]))
    for sweeps in (1, 2):
        assert np.max(np.abs(res[sweeps][0] - res[0][0])) <= 1e-10 * np.max(np.abs(res[0][0]))
        assert abs(res[sweeps][1] - res[0][1]) <= 1e-10 * abs(res[0][1])
        assert res[sweeps][2] < res[0][2]
    model.close()


def test_fp32_refined_forward_gives_the_fp64_gradients(ml, gpu_ctx, wl):
    """The forward solves in the FP32-refined form (then polished in FP64 to
    the rounding floor, as the adjoint requires R(phi^n) = 0): loss and
    gradients equal those of the FP64 forward to 1e-8 relative."""
    torch, adj, case, model, tphi, tchi = _setup(ml, gpu_ctx, wl)
    F = wl.flux_profile(32, F0=10.0)
    L64, g64, _, s64 = _adjoint_grads(torch, adj, model, F, 5.0, tphi, tchi)
    m32 = adj.HeatFluxAdjoint(gpu_ctx, case, checkpoint=5,
                              opts=ml.default_opts(max_iter=400, hang=-1.0, precision=ml.PRECISION_FP32_REFINED))
    L32, g32, gperm, s32 = _adjoint_grads(torch, adj, m32, F, 5.0, tphi, tchi)
    print("FP64 forward: loss %.12e dL/dSt %.12e | FP32-refined forward: loss %.12e dL/dSt %.12e"
          % (L64, s64, L32, s32))
    assert gperm == 0.0
    assert abs(L32 - L64) <= 1e-10 * abs(L64)
    assert np.max(np.abs(g32 - g64)) <= 1e-8 * np.max(np.abs(g64))
    assert abs(s32 - s64) <= 1e-8 * abs(s64)
    m32.close()
    model.close()
