"""The device 3x3 SVD (crHipSvd3: the SVD inside PB-MPM's material constraint)
on its own, against the reference's svd() and against closed forms.

The reference fixture (tests/golden/svd3.npz, written by the reference's CPU
build of include/math/svd3.h) holds the reference's own test matrix, random
matrices, a repeated singular value, the identity, a negative determinant, a
rank-2 matrix and a 1e8 condition number.  The device build contracts the
SVD's arithmetic to FMAs and takes the hardware reciprocal square root, as the
reference's CUDA build does, so it matches the CPU fixture to rounding, not
bit for bit.  Bounds (relative to the matrix's largest singular value):
  * singular values within 2e-5 of the reference's;
  * the factors u and vt within 1e-4 of the reference's where the singular
    values are separated (gap above 1e-2), and their product u vt (the
    rotation PB-MPM uses) within 1e-4 wherever the matrix is far from
    singular;
  * u diag(s) vt = A^T as closely as the oracle's FP64 build of the same
    algorithm (its fixed 4 Jacobi sweeps bound the accuracy) plus 2e-5, u and
    vt orthogonal to 1e-5 with determinant +1, and the sign of det A carried
    by the smallest singular value.
"""
import os

import numpy as np
import pytest

import cpu_checkers as CC

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b, scale):
    return np.abs(a - b).max(axis=tuple(range(1, a.ndim))) / scale


def test_device_svd_matches_the_reference_fixture():
    import crmpm_hip as H

    g = np.load(os.path.join(GOLDEN, "svd3.npz"))
    A = g["A"]
    U, Vt, S = H.svd3(A)
    scale = np.abs(g["S"]).max(axis=1)
    es = _rel(S, g["S"], scale)
    print(f"singular values: max rel err {es.max():.2e}")
    assert es.max() <= 2e-5
    srt = np.sort(np.abs(g["S"]), axis=1)
    gap = np.minimum(srt[:, 1] - srt[:, 0], srt[:, 2] - srt[:, 1]) / scale
    sep = gap > 1e-2
    eu = _rel(U, g["U"], 1.0)
    ev = _rel(Vt, g["Vt"], 1.0)
    print(f"factors (separated singular values, {sep.sum()} of {len(A)}): u {eu[sep].max():.2e} vt {ev[sep].max():.2e}")
    assert eu[sep].max() <= 1e-4 and ev[sep].max() <= 1e-4
    well = srt[:, 0] / scale > 1e-3
    er = _rel(U @ Vt, g["U"] @ g["Vt"], 1.0)
    print(f"rotation u vt ({well.sum()} matrices far from singular): {er[well].max():.2e}")
    assert er[well].max() <= 1e-4
    # the cases the fixture was built for
    assert not sep[-5] and not sep[-4]  # repeated singular values, identity
    assert np.linalg.det(A[-3]) < 0 and srt[-3, 0] > 0  # negative determinant
    assert srt[-2, 0] / scale[-2] < 1e-3  # rank 2
    assert scale[-1] / srt[-1, 0] > 1e7  # condition number 1e8


def test_device_svd_closed_forms():
    """Orthogonal rotations, the sign of det A, and u diag(s) vt = A^T as
    closely as the algorithm allows: its fixed 4 sweeps leave some matrices
    reconstructed to only 2.7e-3 even in FP64 (the oracle's FP64 build shows
    the same), so the device's error is held to the FP64 build's plus
    rounding, per matrix."""
    import crmpm_hip as H

    g = np.load(os.path.join(GOLDEN, "svd3.npz"))
    rng = np.random.default_rng(11)
    A = np.concatenate([g["A"], rng.standard_normal((2000, 3, 3)).astype(np.float32)])
    U, Vt, S = H.svd3(A)
    A64 = A.astype(np.float64)
    scale = np.linalg.svd(A64, compute_uv=False)[:, 0]
    rec = _rel((U.astype(np.float64) * S[:, None, :]) @ Vt, A64.transpose(0, 2, 1), scale)
    U64, Vt64, S64 = CC.svd3("oracle", A, fp64=True)
    rec64 = _rel((U64 * S64[:, None, :]) @ Vt64, A64.transpose(0, 2, 1), scale)
    print(f"reconstruction of A^T: device {rec.max():.2e} (median {np.median(rec):.1e}), FP64 oracle {rec64.max():.2e}")
    assert (rec - rec64).max() <= 2e-5
    I = np.eye(3)
    assert np.abs(U @ U.transpose(0, 2, 1) - I).max() <= 1e-5
    assert np.abs(Vt @ Vt.transpose(0, 2, 1) - I).max() <= 1e-5
    assert np.abs(np.linalg.det(U) - 1).max() <= 1e-5 and np.abs(np.linalg.det(Vt) - 1).max() <= 1e-5
    # a negative determinant shows as a negative singular value
    dA = np.linalg.det(A64)
    clear = np.abs(dA) > 1e-3 * scale ** 3
    assert np.array_equal(np.sign(np.prod(S[clear], axis=1)), np.sign(dA[clear]))
    # singular values agree with LAPACK's up to sign, as closely as the FP64 build's
    ref = np.linalg.svd(A64, compute_uv=False)
    err = np.abs(np.sort(np.abs(S), 1)[:, ::-1] - ref).max(axis=1) / scale
    err64 = np.abs(np.sort(np.abs(S64), 1)[:, ::-1] - ref).max(axis=1) / scale
    assert (err - err64).max() <= 2e-5


def test_device_svd_near_identity_matches_the_oracle():
    """The PB constraint's regime: (I + C dt) F close to a rotation times a
    mild stretch.  Against the FP32 oracle (the reference's evaluation order)."""
    import crmpm_hip as H

    rng = np.random.default_rng(5)
    n = 4096
    q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    q *= np.sign(np.linalg.det(q))[:, None, None]
    A = (q @ (np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3)))).astype(np.float32)
    U, Vt, S = H.svd3(A)
    Uo, Vto, So = CC.svd3("oracle", A)
    es = np.abs(S - So).max()
    er = np.abs(U @ Vt - Uo @ Vto).max()
    print(f"near identity, against the oracle: singular values {es:.2e}, rotation {er:.2e}")
    assert es <= 2e-5 and er <= 2e-5


def test_device_svd_arguments():
    import crmpm_hip as H

    lib = H.load_library()
    err_arg = 1  # CRMPM_HIP_ERR_ARG
    assert lib.crHipSvd3(None, -1, None, None, None, None) == err_arg
    assert lib.crHipSvd3(None, 2, None, None, None, None) == err_arg
    assert lib.crHipSvd3(None, 0, None, None, None, None) == 0
