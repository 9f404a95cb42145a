"""GPU adjoint of the PB-MPM substep (crHipAdjoint* on a PB solver, the SVD's
polar-rotation derivative included) against the FP64 autodiff restatement
(oracle/pb_autodiff.py), which tests/test_pb_adjoint_cpu.py pins to the C++
oracle's FP64 PB trajectories and to central finite differences.

The comparison is a vector-Jacobian product with the GPU's own dL/dx, so it
measures the adjoint, not the FP32 forward's effect on the residual.  Bounds
(relative, 2-norm over particles): no further from the FP64 gradient than
twice what an FP32 evaluation of the same restatement gives (the FP32
states alone move a gradient that much, and the GPU's rounding -- other
summation orders, the 4-sweep SVD -- is another sample of the same size;
a wrong or missing term gives errors of order one; at 20 iterations about 4e-2 for
alpha and the ratio, whose gradient contracts the constraint's small residual
(B F^-T - I) - dt C), and never above 1e-3 for the initial velocity and 1e-1
for alpha and the ratio.  The gradient does not
depend on the checkpoint interval (bitwise: every reverse step re-runs its
substep from the same start state).
"""
import os
import sys

import numpy as np
import pytest
import torch

import cpu_checkers as CC
import crmpm_scenes as S
from test_pb_adjoint_cpu import pb_block, torch_rollout

sys.path.insert(0, os.path.join(CC.ROOT, "oracle"))
import pb_autodiff as PA  # noqa: E402

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp32_rollout(sc, alpha, ratio, v0, steps):
    """The restatement in FP32 (same constants rounded), for the bound."""
    c = PA.consts_from_scene(sc, alpha=alpha, ratio=ratio)
    c = {k: (v.to(torch.float32) if torch.is_tensor(v) and v.dtype == torch.float64 else v) for k, v in c.items()}
    n = len(sc["xm"])
    x0 = torch.as_tensor(sc["xm"][:, :3], dtype=torch.float32)
    F0 = torch.eye(3, dtype=torch.float32).repeat(n, 1, 1)
    C0 = torch.zeros(n, 3, 3, dtype=torch.float32)
    return PA.rollout(x0, v0, F0, C0, c, steps)


def run_case(sc, steps, every, tape_budget=None, scatter_waves=1):
    import crmpm_hip as H

    n = len(sc["xm"])
    sim = H.DifferentiablePb(sc, steps, checkpoint_every=every)
    if tape_budget is not None:
        sim.set_tape_budget(tape_budget)
    sim.solver.set_scatter_waves(scatter_waves)
    try:
        dev = sim.solver.device
        alpha = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
        ratio = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
        v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
        x = sim.forward(alpha, ratio, v0).cpu().numpy().astype(np.float64)
        rng = np.random.default_rng(17)
        gx = rng.normal(0.0, 1.0, (n, 3)) / n
        da, dr, dv = sim.backward(torch.as_tensor(gx, device=dev))
        return x, gx, da.cpu().numpy(), dr.cpu().numpy(), dv.double().cpu().numpy()
    finally:
        sim.close()


def reference_grads(sc, steps, gx, dtype=torch.float64):
    alpha = torch.as_tensor(sc["params"][:, 0], dtype=dtype).clone().requires_grad_(True)
    ratio = torch.as_tensor(sc["params"][:, 1], dtype=dtype).clone().requires_grad_(True)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=dtype).clone().requires_grad_(True)
    if dtype == torch.float64:
        x, _, _, _ = torch_rollout(sc, alpha, ratio, v0, steps=steps)
    else:
        x, _, _, _ = fp32_rollout(sc, alpha, ratio, v0, steps)
    grads = torch.autograd.grad(x, (alpha, ratio, v0), grad_outputs=torch.as_tensor(gx, dtype=dtype), allow_unused=True)
    # without solver iterations alpha and the ratio do not enter the rollout
    ga, gr, gv = (torch.zeros_like(t) if g is None else g for g, t in zip(grads, (alpha, ratio, v0)))
    return x.detach().double().numpy(), ga.double().numpy(), gr.double().numpy(), gv.double().numpy()


CASES = {
    # the CPU test's block: 120 particles, 4 iterations, 8 substeps
    "block120_k4": dict(scene=lambda: pb_block(seed=7), steps=8, every=3),
    # 960 particles, 20 iterations (the headline's count), 12 substeps, a
    # co-rotational block with every 7th particle of a material the
    # constraint leaves alone
    "block960_k20": dict(scene=lambda: pb_block(seed=5, n=(12, 10, 8), iterations=20, material=S.COROTATIONAL),
                         steps=12, every=5, inactive_every=7),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_pb_adjoint_matches_fp64_autograd(case):
    spec = CASES[case]
    sc = spec["scene"]()
    if spec.get("inactive_every"):
        sc["type"][:: spec["inactive_every"]] = S.LINEAR
    n, steps = len(sc["xm"]), spec["steps"]
    print(f"PB adjoint {case}: {n} particles, {sc['iterations']} iterations, {steps} substeps", flush=True)
    x, gx, da, dr, dv = run_case(sc, steps, spec["every"])
    xr, ra, rr, rv = reference_grads(sc, steps, gx)
    _, fa, fr, fv = reference_grads(sc, steps, gx, torch.float32)
    ex = float(np.abs(x - xr).max())
    errs = {"alpha": (relnorm(da, ra), relnorm(fa, ra), 1e-1), "ratio": (relnorm(dr, rr), relnorm(fr, rr), 1e-1),
            "v0": (relnorm(dv, rv), relnorm(fv, rv), 1e-3)}
    print(f"  forward x vs FP64 restatement {ex:.2e} m", flush=True)
    for k, (eg, e32, cap) in errs.items():
        print(f"  d/d{k}: GPU vs FP64 {eg:.3e}, FP32 restatement vs FP64 {e32:.3e}", flush=True)
    assert ex < 1e-5
    for k, (eg, e32, cap) in errs.items():
        assert eg <= min(cap, max(2.0 * e32, 1e-5)), (k, eg, e32)


def test_pb_adjoint_segment_tape_equals_substep_reruns():
    """The backward pass reverses a segment while re-running the one before
    it when two segments' tapes fit the tape budget, tapes one segment at a
    time when one fits, and otherwise re-runs each substep before reversing
    it: the same gradients, bitwise (7 substeps in segments of 3: three
    segments, the last one short)."""
    import crmpm_hip as H

    sc = pb_block(seed=21, iterations=6)
    steps = 7
    probe = H.DifferentiablePb(sc, steps, checkpoint_every=3)
    seg = probe.segment_tape_bytes()
    probe.close()
    assert seg > 0
    runs = [run_case(sc, steps, 3), run_case(sc, steps, 3, tape_budget=seg + seg // 2),
            run_case(sc, steps, 3, tape_budget=0)]
    for r in runs[1:]:
        for u, w in zip(r[2:], runs[0][2:]):
            assert np.array_equal(u, w)


def test_pb_adjoint_checkpoint_interval_independent():
    sc = pb_block(seed=3)
    steps = 9
    a = run_case(sc, steps, 2)
    b = run_case(sc, steps, steps)
    for u, w in zip(a[2:], b[2:]):
        assert np.array_equal(u, w)


def test_pb_adjoint_without_iterations_is_ballistic():
    """No solver iterations: a substep is the integration alone, so the final
    positions are x0 + steps dt v0 + const and dL/dv0 = steps dt dL/dx
    exactly (closed form), with zero alpha and ratio gradients."""
    sc = pb_block(seed=13, iterations=0)
    steps = 6
    x, gx, da, dr, dv = run_case(sc, steps, 4)
    dt = float(np.float32(sc["dt"]))
    assert np.all(da == 0.0) and np.all(dr == 0.0)
    np.testing.assert_allclose(dv, gx * steps * dt, rtol=1e-5, atol=1e-12)
    xr, ra, rr, rv = reference_grads(sc, steps, gx)
    assert float(np.abs(x - xr).max()) < 1e-6
    assert relnorm(dv, rv) < 1e-5


def test_pb_adjoint_masked_particles():
    """Masked particles neither move nor take part (the PB solver honours the
    mask): their v0, alpha and ratio gradients are zero when the loss is on
    positions, and the active particles' gradients equal those of the same
    block without the masked particles (FP64 restatement), at the bounds above."""
    sc = pb_block(seed=5)
    sc["mask"][::5] = 0
    steps = 8
    act = sc["mask"] != 0
    x, gx, da, dr, dv = run_case(sc, steps, 3)
    assert np.all(dv[~act] == 0.0) and np.all(da[~act] == 0.0) and np.all(dr[~act] == 0.0)
    np.testing.assert_array_equal(x[~act], sc["xm"][~act, :3].astype(np.float64))
    sub = {k: (v[act] if isinstance(v, np.ndarray) and v.shape[:1] == act.shape else v) for k, v in sc.items()}
    xr, ra, rr, rv = reference_grads(sub, steps, gx[act])
    _, fa, fr, fv = reference_grads(sub, steps, gx[act], torch.float32)
    assert float(np.abs(x[act] - xr).max()) < 1e-5
    for name, g, r, f, cap in (("alpha", da[act], ra, fa, 1e-1), ("ratio", dr[act], rr, fr, 1e-1),
                               ("v0", dv[act], rv, fv, 1e-3)):
        eg, e32 = relnorm(g, r), relnorm(f, r)
        print(f"  masked block d/d{name}: GPU vs FP64 {eg:.3e}, FP32 restatement vs FP64 {e32:.3e}", flush=True)
        assert eg <= min(cap, max(2.0 * e32, 1e-5)), (name, eg, e32)
