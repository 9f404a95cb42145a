"""The FP64 autodiff restatement of the PB-MPM substep (oracle/pb_autodiff.py),
which checks the GPU's PB adjoint, pinned here on the CPU:
  * its forward trajectory equals the C++ oracle's FP64 PB trajectory (the
    oracle is bit-exact against the reference's CPU PB solver in FP32), to
    the 4-sweep SVD's convergence error;
  * its autograd gradients w.r.t. the constraint parameters (alpha = params0.x,
    ratio = params0.y) and the initial velocity equal central finite
    differences of the C++ oracle in FP64;
  * planted mistakes (the polar rotation of T instead of T^T, the
    untransposed inverse of F) move the forward trajectory far outside the
    agreement.
"""
import os
import sys

import numpy as np
import torch

import cpu_checkers as CC
import crmpm_scenes as S

sys.path.insert(0, os.path.join(CC.ROOT, "oracle"))
import pb_autodiff as PA  # noqa: E402

STEPS = 8


def pb_block(seed=7, n=(6, 5, 4), iterations=4, material=S.NEOHOOKEAN):
    """A small PB block in free flight with a varied velocity field (so the
    constraint works every iteration), per-particle alpha and ratio."""
    rng = S.SplitMix64(seed)
    dx = 1.0 / 32
    h = dx / 2
    N = n[0] * n[1] * n[2]
    dt = 1e-3
    sc = S.new_scene(S.PB, dx, (0, 0, 0), (1, 1, 1), N, dt, iterations=iterations)
    sc["xm"][:, :3] = S.lattice_block(rng, n, (0.45, 0.5, 0.45), (h, h, h), 0.2 * h)
    sc["xm"][:, 3] = 1000.0 * h ** 3
    sc["v"][:, :3] = np.array([0.3, -0.5, 0.1]) + rng.uniform(3 * N, -1.0, 1.0).reshape(N, 3)
    sc["params"][:, 0] = (0.5 / dt) * (0.6 + 0.8 * rng.uniform(N))
    sc["params"][:, 1] = 0.7 + 0.29 * rng.uniform(N)
    sc["type"][:] = material
    return sc


def oracle_final(sc, steps=STEPS):
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    o.step(steps)
    st = o.state()
    o.close()
    return st


def torch_rollout(sc, alpha=None, ratio=None, v0=None, steps=STEPS):
    c = PA.consts_from_scene(sc, alpha=alpha, ratio=ratio)
    n = len(sc["xm"])
    x0 = torch.as_tensor(sc["xm"][:, :3], dtype=torch.float64)
    if v0 is None:
        v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64)
    F0 = torch.eye(3, dtype=torch.float64).repeat(n, 1, 1)
    C0 = torch.zeros(n, 3, 3, dtype=torch.float64)
    return PA.rollout(x0, v0, F0, C0, c, steps)


def forward_errors(sc, steps=STEPS):
    ref = oracle_final(sc, steps)
    x, v, F, C = torch_rollout(sc, steps=steps)
    ex = np.abs(x.numpy() - ref["xm"][:, :3]).max()
    ev = np.abs(v.numpy() - ref["v"]).max() / np.abs(ref["v"]).max()
    # the oracle returns F column-major: element (r, c) at 3 c + r
    eF = np.abs(F.numpy() - ref["F"].reshape(-1, 3, 3).transpose(0, 2, 1)).max()
    return ex, ev, eF


def test_pb_autodiff_forward_matches_oracle():
    for material in (S.NEOHOOKEAN, S.COROTATIONAL):
        sc = pb_block(material=material)
        ex, ev, eF = forward_errors(sc)
        print(f"PB autodiff vs oracle FP64 (type {material}): x {ex:.2e} m, v rel {ev:.2e}, F {eF:.2e}")
        assert ex < 1e-10 and ev < 1e-8 and eF < 1e-10


def test_pb_autodiff_planted_mistakes_are_caught():
    sc = pb_block()
    saved_rot, saved_inv = PA.polar_rotation, PA.inverse_t
    try:
        # the rotation of T rather than of T^T (svd3 factors the transpose)
        PA.polar_rotation = lambda A: saved_rot(A.transpose(1, 2))
        ex, ev, _ = forward_errors(sc)
        print(f"planted polar(T): x {ex:.2e}, v rel {ev:.2e}")
        assert ev > 1e-3
        PA.polar_rotation = saved_rot
        # the inverse of F rather than its transpose (Mat3x3f::inverse)
        PA.inverse_t = lambda F: saved_inv(F).transpose(1, 2)
        ex, ev, _ = forward_errors(sc)
        print(f"planted inverse(F): x {ex:.2e}, v rel {ev:.2e}")
        assert ev > 1e-5
    finally:
        PA.polar_rotation, PA.inverse_t = saved_rot, saved_inv


def loss_of(x, v, tx, tv):
    return ((x - tx) ** 2).sum(1).mean() + 1e-4 * ((v - tv) ** 2).sum(1).mean()


def test_pb_autodiff_gradients_match_finite_differences():
    sc = pb_block(seed=11)
    n = len(sc["xm"])
    ref = oracle_final(sc)
    rng = np.random.default_rng(4)
    tx = torch.as_tensor(ref["xm"][:, :3] + rng.normal(0, 2e-4, (n, 3)), dtype=torch.float64)
    tv = torch.as_tensor(ref["v"] + rng.normal(0, 2e-2, (n, 3)), dtype=torch.float64)
    alpha = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64).clone().requires_grad_(True)
    ratio = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64).clone().requires_grad_(True)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64).clone().requires_grad_(True)
    x, v, _, _ = torch_rollout(sc, alpha, ratio, v0)
    loss_of(x, v, tx, tv).backward()

    def fd_loss(p, col=None, dval=0.0, dv=None):
        s2 = {k: (a.copy() if isinstance(a, np.ndarray) else a) for k, a in sc.items()}
        if col is not None:
            s2["params"][p, col] = np.float32(sc["params"][p, col] + dval)
        if dv is not None:
            s2["v"][p, :3] = (sc["v"][p, :3].astype(np.float64) + dv).astype(np.float32)
        st = oracle_final(s2)
        return float(loss_of(torch.as_tensor(st["xm"][:, :3]), torch.as_tensor(st["v"]), tx, tv))

    for p in (0, n // 2, n - 1):
        for col, grad, h in ((0, alpha.grad, 1.0), (1, ratio.grad, 1e-3)):
            hp = float(np.float32(sc["params"][p, col] + h)) - float(sc["params"][p, col])
            hm = float(np.float32(sc["params"][p, col] - h)) - float(sc["params"][p, col])
            fd = (fd_loss(p, col, h) - fd_loss(p, col, -h)) / (hp - hm)
            print(f"particle {p}: dL/dparams0[{col}] autograd {float(grad[p]):.6e} fd {fd:.6e}")
            assert abs(fd - float(grad[p])) <= 2e-3 * abs(fd)
        hv = 1e-4
        vp = float(np.float32(sc["v"][p, 1] + hv)) - float(sc["v"][p, 1])
        vm = float(np.float32(sc["v"][p, 1] - hv)) - float(sc["v"][p, 1])
        fdv = (fd_loss(p, dv=np.array([0, vp, 0])) - fd_loss(p, dv=np.array([0, vm, 0]))) / (vp - vm)
        print(f"particle {p}: dL/dv0_y autograd {float(v0.grad[p, 1]):.6e} fd {fdv:.6e}")
        assert abs(fdv - float(v0.grad[p, 1])) <= 2e-3 * abs(fdv)
