"""FP64 differentiable restatement of the reference's PB-MPM substep (torch,
CPU, autograd) -- test infrastructure only: the GPU adjoint of the PB substep
is checked against this module's autograd gradients, and this module is
pinned against the C++ oracle's FP64 PB trajectories and against central
finite differences of the C++ oracle.  Nothing in the product imports it.

Follows src/solver/cpu_pb_mpm_solver.cpp:93-106 (K iterations of reset grid,
P2G, grid update, G2P, then the particle integration) with
src/solver/pb_mpm_algorithm.h:
  binning and weights  :173-178 (base = trunc(s - 0.5)), shape_functions.h:44-57
  P2G                  :150-231: solveMaterialConstraints first
                       (constitutive_models.h:123-154), then
                       p_i += w m (v + C d), m_i += w m
  grid update          :233-284 (cutoff 1e-8 keeps raw momentum, wall band of 2
                       nodes; no gravity on the grid)
  G2P                  :286-357 (v = sum w u_i, C = 4/dx sum w outer(w u_i, d)
                       with the reference's outer-product orientation)
  integration          :359-408 (x += dt v, clamp, v += g dt,
                       F = (I + dt C) F)
The constraint: T = (I + dt C) F; R = the rotation of svd3(T), which factors
T^T (include/math/svd3.h:1265-1275), so u vt is the polar rotation of T^T;
Q = T / (sign(det T) sqrt(clamp(|det T|, 0.1, 1000))); B = r R + (1 - r) Q;
C += alpha ((B inverse(F) - I) - dt C), inverse() being the transposed inverse
(include/math/mat33.h:256-283).  The polar rotation is taken here by Newton's
iteration X <- (X + X^-T) / 2 from T^T (Higham), converged to FP64 precision:
the reference's 4-sweep Jacobi SVD reaches the same rotation to its own
convergence error, which near a rotation (every test scene here) is far below
FP32 rounding; det T > 0 is required (the iteration then converges to a
proper rotation, as the SVD's u vt is).
Scenes without rigid shapes only (the adjoint's scope).  Matrices are
[n, row, col].
"""
import itertools

import numpy as np
import torch

OFFS = torch.tensor(list(itertools.product(range(3), repeat=3)), dtype=torch.float64)  # (a, b, c), c fastest
NEOHOOKEAN, COROTATIONAL = 1, 0
NEWTON_ITERATIONS = 30


def polar_rotation(A):
    """Orthogonal polar factor of A (det A > 0): Newton's iteration."""
    X = A
    for _ in range(NEWTON_ITERATIONS):
        X = 0.5 * (X + torch.linalg.inv(X).transpose(1, 2))
    return X


def inverse_t(F):
    """Mat3x3f::inverse(): the transposed inverse, zero where |det F| < 1e-6."""
    d = torch.linalg.det(F)
    ok = (d.abs() >= 1e-6)[:, None, None]
    safe = torch.where(ok, F, torch.eye(3, dtype=F.dtype).expand_as(F))
    return torch.where(ok, torch.linalg.inv(safe).transpose(1, 2), torch.zeros_like(F))


def constraint(C, F, Fi, alpha, ratio, typ, dt):
    """solveMaterialConstraints (constitutive_models.h:123-154); other
    material types leave C as it is."""
    eye = torch.eye(3, dtype=C.dtype).expand_as(C)
    T = (eye + C * dt) @ F
    d = torch.linalg.det(T)
    assert bool((d > 0).all()), "the restatement's polar rotation needs det T > 0"
    R = polar_rotation(T.transpose(1, 2))
    s = torch.sign(d) * torch.sqrt(torch.clamp(d.abs(), 0.1, 1000.0))
    Q = T / s[:, None, None]
    B = ratio[:, None, None] * R + (1.0 - ratio)[:, None, None] * Q
    Cn = C + alpha[:, None, None] * ((B @ Fi - eye) - C * dt)
    act = ((typ == NEOHOOKEAN) | (typ == COROTATIONAL))[:, None, None]
    return torch.where(act, Cn, C)


def substep(state, consts):
    """One PB substep.  state = (x, v, F, C); consts = dict(m, alpha, ratio,
    typ, g, dt, dx, invdx, lo, hi, dims, iterations).  Returns the new state."""
    x, v, F, C = state
    m, alpha, ratio, typ = consts["m"], consts["alpha"], consts["ratio"], consts["typ"]
    dt, dx = consts["dt"], consts["dx"]
    lo, hi, dims = consts["lo"], consts["hi"], consts["dims"]
    g = consts["g"]
    invdx = consts.get("invdx", 1.0 / dx)
    n = x.shape[0]
    # weights: positions do not move during the iterations
    s = (x - lo) * invdx
    base = torch.trunc(s.detach() - 0.5)
    fx = s - base
    W = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], 1)  # [n, i, axis]
    offs = OFFS.to(x.dtype)
    oi = OFFS.long()
    wt = W[:, oi[:, 0], 0] * W[:, oi[:, 1], 1] * W[:, oi[:, 2], 2]  # [n, 27]
    dist = (offs[None] - fx[:, None, :]) * dx  # [n, 27, 3]
    node = (base[:, None, :] + offs[None]).long()
    inside = ((node >= 0) & (node < dims[None, None, :])).all(-1)
    nx, ny, nz = (int(d) for d in dims)
    flat = node[..., 0] + nx * (node[..., 1] + ny * node[..., 2])
    flat = torch.where(inside, flat, torch.zeros_like(flat))
    wt_in = wt * inside
    nodes = nx * ny * nz
    M = torch.zeros(nodes, dtype=x.dtype).index_add(0, flat.reshape(-1), (wt_in * m[:, None]).reshape(-1))
    heavy = M >= 1e-8
    idx = torch.arange(nodes)
    ii, jj, kk = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    band = torch.stack([(ii < 2) | (ii > nx - 3), (jj < 2) | (jj > ny - 3), (kk < 2) | (kk > nz - 3)], 1)
    Fi = inverse_t(F)
    for _ in range(consts["iterations"]):
        Cc = constraint(C, F, Fi, alpha, ratio, typ, dt)
        contrib = (m[:, None, None] * (v[:, None, :] + torch.einsum("nrc,nkc->nkr", Cc, dist))) * wt_in[..., None]
        P = torch.zeros(nodes, 3, dtype=x.dtype).index_add(0, flat.reshape(-1), contrib.reshape(-1, 3))
        U = torch.where(heavy[:, None], P / torch.where(heavy, M, torch.ones_like(M))[:, None], P)
        U = torch.where(heavy[:, None] & band, torch.zeros_like(U), U)
        Ui = U[flat] * inside[..., None]
        v = (Ui * wt[..., None]).sum(1)
        C = 4.0 * invdx * torch.einsum("nk,nkr,nkc->nrc", wt, dist, Ui)
    xnew = torch.minimum(torch.maximum(x + dt * v, lo), hi)
    vnew = v + g * dt
    eye = torch.eye(3, dtype=x.dtype).expand(n, 3, 3)
    Fnew = (eye + dt * C) @ F
    return xnew, vnew, Fnew, C


def rollout(x0, v0, F0, C0, consts, steps):
    state = (x0, v0, F0, C0)
    for _ in range(steps):
        state = substep(state, consts)
    return state


def consts_from_scene(scene, alpha=None, ratio=None):
    """Constant inputs of a PB scene (masses, constraint parameters, grid) in
    FP64; alpha and ratio may be passed as tensors to differentiate them."""
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
    ext = np.asarray(scene["hi"], np.float32) - np.asarray(scene["lo"], np.float32)
    dims = [int(np.float32(e) / np.float32(scene["dx"]) + np.float32(1.0)) for e in ext]
    # the solver interfaces take dt, dx and gravity as FP32 (include/solver/mpm_solver.h:63-190)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    return dict(m=t(scene["xm"][:, 3]), alpha=t(scene["params"][:, 0]) if alpha is None else alpha,
                ratio=t(scene["params"][:, 1]) if ratio is None else ratio, typ=torch.as_tensor(scene["type"]),
                g=t(f32(scene["gravity"])), dt=float(f32(scene["dt"])), dx=float(f32(scene["dx"])),
                invdx=float(np.float32(1.0) / np.float32(scene["dx"])),
                lo=t(scene["lo"]), hi=t(scene["hi"]), dims=torch.tensor(dims), iterations=int(scene["iterations"]))
