"""GPU adjoint of the MLS substep (crHipAdjoint*) against the FP64 autodiff
restatement, which tests/test_adjoint_cpu.py pins to the C++ oracle's FP64
trajectories and to central finite differences.

Tolerances (relative, in the 2-norm over particles, 30 substeps):
  * v0: within 1e-4 of the FP64 gradient;
  * log E: no further from the FP64 gradient than 1.5 times what an FP32
    evaluation of the same restatement gives (about 1e-2 in this strongly
    deforming test block, where the log E gradient is a small difference of
    stress terms; FP32-rounded states alone, with FP64 arithmetic, leave
    about 4e-3), and within 2e-2;
  * the gradient does not depend on the checkpoint interval (bitwise).
The differentiable block at test size (4,096 particles of the 65,536-particle
GRF block's 96^3 grid, 100 substeps, checkpoints every 10, targets from a
hidden field plus 1 mm noise) is checked against the FP64 autograd
restatement with the same dL/dx (log E no further than 1.5 times an FP32
evaluation of the restatement and at most 1e-2; v0 within 1e-4) and against the
FP64 oracle's central differences on 32 particles (tests/golden/
adjoint_block4096_fd.npz): 1e-2 in log E, 1e-3 in v0, where the FP32
restatement itself is off by 3.6e-3 and 1.5e-4.
"""
import os
import sys

import numpy as np
import pytest
import torch

import cpu_checkers as CC

import crmpm_scenes as S
sys.path.insert(0, os.path.join(CC.ROOT, "oracle"))
import mls_autodiff as AD  # noqa: E402
from test_adjoint_cpu import STEPS, block_scene, fd_directional, torch_rollout

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_grads(sim, lam, mu, v0, target):
    x = sim.forward(lam, mu, v0)
    n = x.shape[0]
    gx = 2.0 * (x.double() - target) / n
    dlam, dmu, dv0 = sim.backward(gx)
    g_loge = (dlam * lam + dmu * mu).cpu().numpy()
    return x.cpu().numpy().astype(np.float64), g_loge, dv0.double().cpu().numpy()


@pytest.mark.parametrize("material", [S.COROTATIONAL, S.NEOHOOKEAN])
def test_adjoint_matches_fp64_autograd(material):
    import crmpm_hip as H

    sc = block_scene(seed=9, material=material)
    n = len(sc["xm"])
    print(f"adjoint material {material}: {n} particles, {STEPS} substeps", flush=True)
    sim = H.DifferentiableMls(sc, STEPS, checkpoint_every=7)
    dev = sim.solver.device
    lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
    mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
    V0 = sim.solver.state()["V0"]
    rng = np.random.default_rng(3)
    # reference forward in FP64 on the CPU; the loss is MSE to perturbed final positions
    logs = torch.zeros(n, dtype=torch.float64, requires_grad=True)
    v0c = v0.cpu().clone().requires_grad_(True)
    xr, _, _, _ = torch_rollout(sc, V0, lam.cpu() * torch.exp(logs), mu.cpu() * torch.exp(logs), v0c)
    target = xr.detach() + torch.as_tensor(rng.normal(0, 1e-3, (n, 3)))
    x, g_loge, g_v0 = gpu_grads(sim, lam, mu, v0, target.to(dev))
    # FP64 vector-Jacobian product with the GPU's own dL/dx, so the comparison
    # measures the adjoint, not the FP32 forward's effect on the residual
    gx = torch.as_tensor(2.0 * (x - target.numpy()) / n)
    ref_loge, ref_v0 = torch.autograd.grad(xr, (logs, v0c), grad_outputs=gx)
    ex = float(np.abs(x - xr.detach().numpy()).max())
    el, ev = relnorm(g_loge, ref_loge.numpy()), relnorm(g_v0, ref_v0.numpy())
    # the same restatement evaluated in FP32, same cotangent
    c32 = AD.consts_to(AD.consts_from_scene(sc, V0), torch.float32)
    l32 = torch.zeros(n, dtype=torch.float32, requires_grad=True)
    c32["lam"] = c32["lam"] * torch.exp(l32)
    c32["mu"] = c32["mu"] * torch.exp(l32)
    v32 = v0.cpu().float().clone().requires_grad_(True)
    x32, _, _, _ = AD.rollout(torch.as_tensor(sc["xm"][:, :3], dtype=torch.float32), v32,
                              torch.eye(3).repeat(n, 1, 1), torch.zeros(n, 3, 3), c32, STEPS)
    g32 = torch.autograd.grad(x32, (l32,), grad_outputs=gx.float())[0].double().numpy()
    e32 = relnorm(g32, ref_loge.numpy())
    print(f"  forward max |x - x64| {ex:.2e} m; gradient rel err: logE {el:.2e} (FP32 restatement {e32:.2e}) "
          f"v0 {ev:.2e}", flush=True)
    assert ex < 1e-6
    assert el <= 1.5 * e32 and el < 2e-2
    assert ev < 1e-4
    sim.close()


def test_checkpoint_interval_does_not_change_gradients():
    import crmpm_hip as H

    sc = block_scene(seed=4)
    n = len(sc["xm"])
    out = []
    for every in (1, 7, STEPS):
        print(f"checkpoint every {every}", flush=True)
        sim = H.DifferentiableMls(sc, STEPS, checkpoint_every=every)
        dev = sim.solver.device
        lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
        mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
        v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
        target = torch.full((n, 3), 0.5, dtype=torch.float64, device=dev)
        out.append(gpu_grads(sim, lam, mu, v0, target))
        sim.close()
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.array_equal(a, b)


def test_overlapped_reverse_equals_sequential():
    """With a second set of segment buffers within the tape budget (the
    default) each segment's re-run overlaps the next segment's reverse on
    another stream; with none the segments run one after the other.  The
    gradients are bitwise the same, backward after backward."""
    import crmpm_hip as H

    sc = block_scene(seed=8)
    n = len(sc["xm"])
    out = []
    for budget in (None, 0):
        sim = H.DifferentiableMls(sc, STEPS, checkpoint_every=4)
        if budget is not None:
            sim.set_tape_budget(budget)
        dev = sim.solver.device
        lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
        mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
        v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
        target = torch.full((n, 3), 0.5, dtype=torch.float64, device=dev)
        out.append([gpu_grads(sim, lam, mu, v0, target) for _ in range(2)])
        sim.close()
    for a, b in zip(out[0] + out[0][:1], out[1] + out[0][1:]):
        for u, w in zip(a, b):
            assert np.array_equal(u, w)


def test_rollout_is_a_torch_autograd_function():
    import crmpm_hip as H

    sc = block_scene(seed=11)
    n = len(sc["xm"])
    sim = H.DifferentiableMls(sc, STEPS, checkpoint_every=10)
    dev = sim.solver.device
    E = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev) * 2 * 1.3  # mu = E / (2 (1 + nu))
    logE = torch.log(E).clone().requires_grad_(True)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev).clone().requires_grad_(True)
    nu = 0.3
    lam = torch.exp(logE) * nu / ((1 + nu) * (1 - 2 * nu))
    mu = torch.exp(logE) / (2 * (1 + nu))
    x = sim.rollout(lam, mu, v0)
    target = torch.full((n, 3), 0.5, dtype=torch.float64, device=dev)
    loss = ((x.double() - target) ** 2).sum(1).mean()
    loss.backward()
    _, g_loge, g_v0 = gpu_grads(sim, lam.detach(), mu.detach(), v0.detach(), target)
    print(f"autograd.Function: |dL/dlogE| {np.abs(g_loge).max():.3e}", flush=True)
    assert relnorm(logE.grad.cpu().numpy(), g_loge) < 1e-12
    assert relnorm(v0.grad.cpu().numpy(), g_v0) < 1e-12
    sim.close()


def test_block_gradients_against_fp64_autograd_and_central_differences():
    import adjoint_block as AB
    import crmpm_hip as H

    fx = np.load(os.path.join(CC.ROOT, "tests", "golden", "adjoint_block4096_fd.npz"))
    sc, _ = AB.scene()
    n = len(sc["xm"])
    print(f"adjoint block: {n} particles, {AB.STEPS} substeps, checkpoints every {AB.EVERY}", flush=True)
    sim = H.DifferentiableMls(sc, AB.STEPS, checkpoint_every=AB.EVERY)
    dev = sim.solver.device
    lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
    mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
    V0 = sim.solver.state()["V0"]
    target = torch.as_tensor(fx["target"], dtype=torch.float64)
    x = sim.forward(lam, mu, v0).double()
    gx_dev = 2.0 * (x - target.to(dev)) / n
    dlam, dmu, dv0 = sim.backward(gx_dev)
    sim.close()
    g_lam, g_mu, g_v = dlam.cpu().numpy(), dmu.cpu().numpy(), dv0.double().cpu().numpy()
    lam_c, mu_c = lam.cpu().numpy(), mu.cpu().numpy()
    g_loge = g_lam * lam_c + g_mu * mu_c
    loss = float(((x.cpu() - target) ** 2).sum(1).mean())
    # FP64 and FP32 vector-Jacobian products of the restatement with the GPU's dL/dx
    gx = gx_dev.cpu()

    def vjp(dtype):
        c = AD.consts_from_scene(sc, V0)
        if dtype != torch.float64:
            c = AD.consts_to(c, dtype)
        ls = torch.zeros(n, dtype=dtype, requires_grad=True)
        c["lam"] = c["lam"] * torch.exp(ls)
        c["mu"] = c["mu"] * torch.exp(ls)
        vv = v0.cpu().to(dtype).clone().requires_grad_(True)
        xr, _, _, _ = AD.rollout(torch.as_tensor(sc["xm"][:, :3], dtype=dtype), vv,
                                 torch.eye(3, dtype=dtype).repeat(n, 1, 1), torch.zeros(n, 3, 3, dtype=dtype), c,
                                 AB.STEPS)
        gl, gv = torch.autograd.grad(xr, (ls, vv), grad_outputs=gx.to(dtype))
        return xr.detach().double().numpy(), gl.double().numpy(), gv.double().numpy()

    x64, r_loge, r_v = vjp(torch.float64)
    _, r32_loge, r32_v = vjp(torch.float32)
    ex = float(np.abs(x.cpu().numpy() - x64).max())
    el, ev = relnorm(g_loge, r_loge), relnorm(g_v, r_v)
    e32l, e32v = relnorm(r32_loge, r_loge), relnorm(r32_v, r_v)
    pe, de, pv, dv = fd_directional(fx, g_lam, g_mu, g_v)
    fe, fv = relnorm(pe, de), relnorm(pv, dv)
    print(f"  loss {loss:.4e} (FP64 oracle {float(fx['L0']):.4e}); forward max |x - x64| {ex:.2e} m", flush=True)
    print(f"  vs FP64 autograd: log E {el:.2e} (FP32 restatement {e32l:.2e}), v0 {ev:.2e} (FP32 {e32v:.2e})",
          flush=True)
    print(f"  vs FP64 central differences on 32 particles: log E {fe:.2e}, v0 {fv:.2e}", flush=True)
    assert ex < 1e-6
    assert abs(loss - float(fx["L0"])) <= 1e-3 * float(fx["L0"])
    assert el <= max(1.5 * e32l, 1e-4) and el < 1e-2, (el, e32l)
    assert ev < 1e-4, (ev, e32v)
    assert fe < 1e-2 and fv < 1e-3, (fe, fv)
