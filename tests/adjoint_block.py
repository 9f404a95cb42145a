"""The differentiable MLS block at test size, shared by the adjoint tests and
their fixture generator (tests/golden/make_adjoint_fd.py).

The 96^3-cell grid (dx = 1/96) and the Gaussian-random-field log E of the
65,536-particle block (crmpm_scenes.config4), with 16^3 = 4,096 particles at
the same spacing, 100 substeps of dt = 1e-4, checkpoints every 10.  The
inverse problem's targets are the final positions of the same block under a
hidden log E field (another field seed), plus seeded Gaussian measurement
noise of 1 mm per coordinate, so the residuals x - x* sit far above FP32
resolution of the positions; the loss is mean_p |x_p - x*_p|^2.
"""
import copy

import numpy as np

import cpu_checkers as CC
import crmpm_scenes as S

N_SIDE = 16
STEPS = 100
EVERY = 10
HIDDEN_FIELD_SEED = S.HIDDEN_FIELD_SEED


def scene():
    sc, log_e = S.config4(nx=N_SIDE, ny=N_SIDE, nz=N_SIDE, steps=STEPS)
    return sc, log_e


def hidden_scene():
    return S.config4(nx=N_SIDE, ny=N_SIDE, nz=N_SIDE, steps=STEPS, field_seed=HIDDEN_FIELD_SEED)


def measurement_noise(n):
    return S.gaussian_noise(n, S.TARGET_NOISE_SEED, S.TARGET_NOISE)


def oracle_final_positions(sc, steps=STEPS):
    """Final positions of the FP64 oracle (V0 from computeInitialData on the
    initial positions)."""
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle", fp64=True)
    o.compute_initial_data()
    o.step(steps)
    x = o.state()["xm"][:, :3].copy()
    o.close()
    return x


def targets():
    hid, _ = hidden_scene()
    x = oracle_final_positions(hid)
    return x + measurement_noise(len(x))


def loss(x, target):
    return float(((np.asarray(x, np.float64) - target) ** 2).sum(1).mean())
