"""Fixed parity bounds of the GPU solver against the FP32 oracle, shared by
the GPU parity tests and the CPU test that shows a planted error exceeds them.

Relative error of a field = max |a - b| over the oracle field's max |b|.
  * TOL_10: positions, velocities and F after 10 substeps (and the other
    short runs);
  * TOL_100: after 100 or more substeps.
Measured GPU-against-oracle errors are 3e-8 to 3e-5 (summation order of the
tile sums against the oracle's particle order, the SVD's FMAs and hardware
reciprocal square root, logf); a friction coefficient off by 10% moves the
velocities by 2.5e-3 to 4.7e-3 after 10 substeps.
"""
import numpy as np

TOL_10 = 5e-5
TOL_100 = 1e-4


def rel(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a - b).max()) / scale
