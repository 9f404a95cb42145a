"""The FP64 autodiff restatement (oracle/mls_autodiff.py) that checks the GPU
adjoint, pinned here on the CPU:
  * its forward trajectory equals the C++ oracle's FP64 trajectory (which is
    bit-exact against the reference's CPU solver in FP32);
  * its autograd gradients w.r.t. log E and v0 equal central finite
    differences of the C++ oracle in FP64 (h = 1e-2 in log E, whose O(h^2)
    truncation is ~1e-4 relative; 1e-4 m/s in v0).
A planted sign error in the stress shows up in the gradient comparison.
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

STEPS = 30


def block_scene(seed=5, n=(6, 5, 4), material=S.COROTATIONAL):
    rng = S.SplitMix64(seed)
    dx = 1.0 / 32
    h = dx / 2
    N = n[0] * n[1] * n[2]
    sc = S.new_scene(S.MLS, dx, (0, 0, 0), (1, 1, 1), N, 1e-4)
    pts = S.lattice_block(rng, n, (0.45, 0.5, 0.45), (h, h, h), 0.2 * h)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * h ** 3
    # strongly varying initial velocities, so the stress (and E) matter within a few substeps
    sc["v"][:, :3] = np.array([0.3, -0.5, 0.1]) + rng.uniform(3 * N, -1.0, 1.0).reshape(N, 3)
    logE = np.log(5e3) + rng.uniform(N) * (np.log(5e4) - np.log(5e3))
    lam, mu = S.lame(np.exp(logE), 0.3)
    sc["params"][:, 0] = lam
    sc["params"][:, 1] = mu
    sc["type"][:] = material
    return sc


def oracle_final(sc, steps=STEPS):
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    o.step(steps)
    st = o.state()
    o.close()
    return st


def initial_volume(sc):
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    V0 = o.state()["V0"]
    o.close()
    return V0


def torch_rollout(sc, V0, lam, mu, v0, steps=STEPS):
    c = AD.consts_from_scene(sc, V0, lam=lam, mu=mu)
    n = len(sc["xm"])
    x0 = torch.as_tensor(sc["xm"][:, :3], dtype=torch.float64)
    F0 = torch.eye(3, dtype=torch.float64).repeat(n, 1, 1)
    C0 = torch.zeros(n, 3, 3, dtype=torch.float64)
    return AD.rollout(x0, v0, F0, C0, c, steps)


@pytest.mark.parametrize("material", [S.COROTATIONAL, S.NEOHOOKEAN])
def test_autodiff_forward_matches_oracle(material):
    sc = block_scene(material=material)
    V0 = initial_volume(sc)
    ref = oracle_final(sc)
    lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64)
    mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64)
    x, v, F, C = torch_rollout(sc, V0, lam, mu, v0)
    ex = np.abs(x.numpy() - ref["xm"][:, :3]).max()
    ev = np.abs(v.numpy() - ref["v"]).max() / np.abs(ref["v"]).max()
    print(f"autodiff vs oracle FP64: x {ex:.2e} m, v rel {ev:.2e}")
    assert ex < 1e-12 and ev < 1e-10


def loss_of(x, target):
    return ((x - target) ** 2).sum(1).mean()


def test_autodiff_gradient_matches_finite_differences():
    sc = block_scene(seed=9)
    V0 = initial_volume(sc)
    n = len(sc["xm"])
    rng = np.random.default_rng(3)
    target = torch.as_tensor(oracle_final(sc)["xm"][:, :3] + rng.normal(0, 2e-4, (n, 3)), dtype=torch.float64)
    lam0 = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64)
    mu0 = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64)
    logs = torch.zeros(n, dtype=torch.float64, requires_grad=True)  # log E offset per particle
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64).clone().requires_grad_(True)
    x, _, _, _ = torch_rollout(sc, V0, lam0 * torch.exp(logs), mu0 * torch.exp(logs), v0)
    loss_of(x, target).backward()
    g_loge, g_v0 = logs.grad.numpy(), v0.grad.numpy()

    def fd_loss(p, dlog=0.0, dv=None):
        s2 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
        s2["params"][p, :2] = (sc["params"][p, :2].astype(np.float64) * np.exp(dlog)).astype(np.float32)
        if dv is not None:
            s2["v"][p, :3] = (sc["v"][p, :3].astype(np.float64) + dv).astype(np.float32)
        xT = oracle_final(s2)["xm"][:, :3]
        return float(((xT - target.numpy()) ** 2).sum(1).mean())

    for p in (0, n // 3, n - 1):
        h = 1e-2
        # the float32 parameter storage rounds the step; use the realised step
        lp = np.log(np.float32(sc["params"][p, 0] * np.exp(h)) / sc["params"][p, 0])
        lm = np.log(np.float32(sc["params"][p, 0] * np.exp(-h)) / sc["params"][p, 0])
        fd = (fd_loss(p, h) - fd_loss(p, -h)) / (lp - lm)
        print(f"particle {p}: dL/dlogE autograd {g_loge[p]:.6e} fd {fd:.6e}")
        assert abs(fd - g_loge[p]) <= 2e-3 * abs(fd)
        hv = 1e-4
        vp = float(np.float32(sc["v"][p, 1] + hv)) - float(sc["v"][p, 1])
        vm = float(np.float32(sc["v"][p, 1] - hv)) - float(sc["v"][p, 1])
        fdv = (fd_loss(p, dv=np.array([0, vp, 0])) - fd_loss(p, dv=np.array([0, vm, 0]))) / (vp - vm)
        print(f"particle {p}: dL/dv0_y autograd {g_v0[p, 1]:.6e} fd {fdv:.6e}")
        assert abs(fdv - g_v0[p, 1]) <= 2e-3 * abs(fdv)


def block_gradients_fp64(sc, V0, target, dtype=torch.float64):
    """Loss and dL/dlambda, dL/dmu, dL/dv0 of the differentiable block through
    the autodiff restatement (FP64, or FP32 to measure FP32's own effect)."""
    import adjoint_block as AB

    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), dtype=dtype)  # noqa: E731
    lam = t(sc["params"][:, 0]).requires_grad_(True)
    mu = t(sc["params"][:, 1]).requires_grad_(True)
    v0 = t(sc["v"][:, :3]).requires_grad_(True)
    c = AD.consts_from_scene(sc, V0, lam=lam, mu=mu)
    if dtype != torch.float64:
        c = AD.consts_to(c, dtype)
        c["lam"], c["mu"] = lam, mu
    n = len(sc["xm"])
    x, _, _, _ = AD.rollout(t(sc["xm"][:, :3]), v0, torch.eye(3, dtype=dtype).repeat(n, 1, 1),
                            torch.zeros(n, 3, 3, dtype=dtype), c, AB.STEPS)
    L = ((x - t(target)) ** 2).sum(1).mean()
    L.backward()
    return float(L.detach()), lam.grad.double().numpy(), mu.grad.double().numpy(), v0.grad.double().numpy()


def fd_directional(fx, g_lam, g_mu, g_v):
    """(predicted, finite-difference) changes of the loss over the fixture's
    steps: log E steps as realised in FP32 lambda and mu, and v0 steps."""
    idx = fx["idx"]
    dl = fx["lam_pm"][:, 0].astype(np.float64) - fx["lam_pm"][:, 1]
    dm = fx["mu_pm"][:, 0].astype(np.float64) - fx["mu_pm"][:, 1]
    dv = fx["v_pm"][:, 0].astype(np.float64) - fx["v_pm"][:, 1]
    pred_e = g_lam[idx] * dl + g_mu[idx] * dm
    pred_v = (g_v[idx] * dv).sum(1)
    return pred_e, fx["L_loge"][:, 0] - fx["L_loge"][:, 1], pred_v, fx["L_v"][:, 0] - fx["L_v"][:, 1]


def test_autodiff_block_gradients_match_fd_fixture():
    """The differentiable block at test size (4,096 particles, 96^3 grid, GRF
    log E, 100 substeps, targets from a hidden field plus 1 mm noise): the
    restatement's FP64 loss equals the oracle's, and its gradients predict the
    oracle's FP64 central differences (h = 1e-3 in log E, 1e-4 m/s in v0) on
    32 particles: v0 to 1e-5; log E to 2e-4, the FP64 rounding of the loss
    (about 1e-20 against loss changes of 1e-17 to 2e-15 per step) being the
    limit there, not the O(h^2) truncation."""
    import adjoint_block as AB

    fx = np.load(os.path.join(CC.ROOT, "tests", "golden", "adjoint_block4096_fd.npz"))
    sc, _ = AB.scene()
    V0 = initial_volume(sc)
    L, g_lam, g_mu, g_v = block_gradients_fp64(sc, V0, fx["target"])
    assert abs(L - float(fx["L0"])) <= 1e-10 * float(fx["L0"]), (L, float(fx["L0"]))
    pe, de, pv, dv = fd_directional(fx, g_lam, g_mu, g_v)
    ee = np.linalg.norm(pe - de) / np.linalg.norm(de)
    ev = np.linalg.norm(pv - dv) / np.linalg.norm(dv)
    print(f"loss {L:.6e}; autograd vs central differences: log E {ee:.2e}, v0 {ev:.2e}")
    assert ee <= 2e-4 and ev <= 1e-5, (ee, ev)
    # the steps move the loss far above its FP64 rounding (eps L)
    eps_l = np.finfo(np.float64).eps * L
    assert np.median(np.abs(de)) > 1e4 * eps_l and np.abs(dv).min() > 1e4 * eps_l


def test_planted_stress_sign_error_is_caught():
    """Flipping the sign of the stress term in the restatement breaks the trajectory check."""
    sc = block_scene()
    V0 = initial_volume(sc)
    ref = oracle_final(sc)
    orig = AD.stress
    try:
        AD.stress = lambda F, lam, mu, typ: -orig(F, lam, mu, typ)
        lam = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64)
        mu = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64)
        v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64)
        x, _, _, _ = torch_rollout(sc, V0, lam, mu, v0)
    finally:
        AD.stress = orig
    assert np.abs(x.numpy() - ref["xm"][:, :3]).max() > 1e-9


class _FakeSim:
    """Stands in for DifferentiableMls on the CPU: forward returns x = v0 and
    bumps the generation as the real forward does."""

    def __init__(self):
        self.generation = 0

    def forward(self, lam, mu, v0):
        self.generation += 1
        return v0.detach().clone().float()

    def backward(self, grad_x):
        n = grad_x.shape[0]
        return torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float64), grad_x.float()


def test_rollout_backward_refuses_a_stale_graph():
    """A rollout's backward after another forward on the same sim raises
    instead of returning the later forward's gradients."""
    import crmpm_hip as H

    sim = _FakeSim()
    lam = torch.ones(4, dtype=torch.float64, requires_grad=True)
    v0 = torch.ones(4, 3, dtype=torch.float64, requires_grad=True)
    x1 = H._MlsRollout.apply(lam, lam, v0, sim)
    x1.sum().backward()  # the latest rollout: fine
    assert torch.equal(v0.grad, torch.ones(4, 3, dtype=torch.float64))
    x2 = H._MlsRollout.apply(lam, lam, v0, sim)
    sim.forward(lam, lam, v0)  # e.g. an evaluation forward in between
    with pytest.raises(RuntimeError, match="another forward"):
        x2.sum().backward()
