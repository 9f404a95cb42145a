"""Write tests/golden/adjoint_block4096_fd.npz: central finite differences of
the differentiable block's loss (tests/adjoint_block.py) through the FP64
CPU oracle, for 32 seeded particles, against which the FP64 autodiff
restatement (CPU test) and the GPU adjoint (GPU test) are checked.

Per particle p:
  * log E: lambda_p and mu_p scaled by exp(+-h), h = 1e-3, as the FP32 values
    the solver takes; the fixture keeps those values, so a test compares
    dL/dlambda (lambda+ - lambda-) + dL/dmu (mu+ - mu-) with L+ - L-;
  * v0: a seeded unit direction d, v0 +- 1e-4 m/s d rounded to FP32; a test
    compares dL/dv0 . (v+ - v-) with L+ - L-.
Run from the repository root:  python tests/golden/make_adjoint_fd.py
"""
import copy
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "cressim-mpm_amd", "python"))

import adjoint_block as AB  # noqa: E402
import crmpm_scenes as S  # noqa: E402

H_LOGE = 1e-3
H_V = 1e-4
N_PICK = 32
OUT = os.path.join(HERE, "adjoint_block4096_fd.npz")

_SC = None
_TARGET = None


def _init(sc, target):
    global _SC, _TARGET
    _SC, _TARGET = sc, target


def _loss_with(job):
    p, lam, mu, v = job
    s2 = copy.deepcopy(_SC)
    if lam is not None:
        s2["params"][p, 0] = lam
        s2["params"][p, 1] = mu
    if v is not None:
        s2["v"][p, :3] = v
    return AB.loss(AB.oracle_final_positions(s2), _TARGET)


def main():
    sc, _ = AB.scene()
    n = len(sc["xm"])
    target = AB.targets()
    rng = S.SplitMix64(20261203 + 46)
    idx = np.sort(np.unique((rng.uniform(4 * N_PICK) * n).astype(np.int64)))[:N_PICK]
    assert len(idx) == N_PICK
    dirs = rng.uniform(3 * N_PICK, -1.0, 1.0).reshape(N_PICK, 3)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    f32 = np.float32
    lam_pm = np.zeros((N_PICK, 2), np.float32)
    mu_pm = np.zeros((N_PICK, 2), np.float32)
    v_pm = np.zeros((N_PICK, 2, 3), np.float32)
    jobs = [(-1, None, None, None)]
    for j, p in enumerate(idx):
        lam, mu = float(sc["params"][p, 0]), float(sc["params"][p, 1])
        v = sc["v"][p, :3].astype(np.float64)
        for s, sign in enumerate((1.0, -1.0)):
            lam_pm[j, s] = f32(lam * np.exp(sign * H_LOGE))
            mu_pm[j, s] = f32(mu * np.exp(sign * H_LOGE))
            v_pm[j, s] = (v + sign * H_V * dirs[j]).astype(np.float32)
            jobs.append((int(p), lam_pm[j, s], mu_pm[j, s], None))
            jobs.append((int(p), None, None, v_pm[j, s]))
    with ProcessPoolExecutor(8, initializer=_init, initargs=(sc, target)) as ex:
        losses = list(ex.map(_loss_with, jobs, chunksize=4))
    L0 = losses[0]
    L = np.array(losses[1:]).reshape(N_PICK, 2, 2)  # particle, sign, (log E, v0)
    np.savez_compressed(OUT, target=target, idx=idx, lam_pm=lam_pm, mu_pm=mu_pm, v_pm=v_pm,
                        L_loge=L[:, :, 0], L_v=L[:, :, 1], L0=np.float64(L0), steps=np.int32(AB.STEPS),
                        h_loge=np.float64(H_LOGE), h_v=np.float64(H_V))
    print("wrote", OUT, "loss", L0)


if __name__ == "__main__":
    main()
