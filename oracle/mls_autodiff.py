"""FP64 differentiable restatement of the reference's MLS-MPM substep (torch,
CPU, autograd) -- test infrastructure only: the GPU adjoint's gradients are
checked against this module's autograd gradients, and this module is pinned
against the C++ oracle's FP64 trajectories and against central finite
differences of the C++ oracle.  Nothing in the product imports it.

Follows src/solver/mls_mpm_algorithm.h (the reference's MLS substep):
  binning and weights  :172-174 (base = trunc(s - 0.5)), shape_functions.h:44-57
  P2G                  :147-241 (v += g dt written back, affine = -4 dt V0 tau / dx^2 + m C)
  grid update          :243-294 (cutoff 1e-4 keeps raw momentum, wall band of 2 nodes)
  G2P                  :296-390 (C = 4/dx sum w outer(w v_i, d) with the reference's
                                 outer-product orientation, clamp, F = (I + dt C) F)
  stress               constitutive_models.h:49-75 (StVK for co-rotational, Neo-Hookean)
Scenes without rigid shapes only (the adjoint's scope).  Matrices are [n, row, col].
"""
import itertools

import numpy as np
import torch

OFFS = torch.tensor(list(itertools.product(range(3), repeat=3)), dtype=torch.float64)  # (a, b, c), c fastest


def stress(F, lam, mu, typ):
    """Kirchhoff stress (constitutive_models.h:95-118): co-rotational and linear
    elastic evaluate StVK, Neo-Hookean with J clamped at 1e-6."""
    eye = torch.eye(3, dtype=F.dtype).expand_as(F)
    Ft = F.transpose(1, 2)
    E = 0.5 * (Ft @ F - eye)
    trE = E.diagonal(dim1=1, dim2=2).sum(-1)
    S = lam[:, None, None] * trE[:, None, None] * eye + 2.0 * mu[:, None, None] * E
    t_stvk = F @ S @ Ft
    J = torch.clamp(torch.linalg.det(F), min=1e-6)
    t_nh = mu[:, None, None] * (F @ Ft - eye) + (lam * torch.log(J))[:, None, None] * eye
    nh = (typ == 1)[:, None, None]
    return torch.where(nh, t_nh, t_stvk)


def substep(state, consts):
    """One MLS substep.  state = (x, v, F, C); consts = dict(m, V0, lam, mu, typ,
    g, dt, dx, lo, hi, dims).  Returns the new state."""
    x, v, F, C = state
    m, V0, lam, mu, typ = consts["m"], consts["V0"], consts["lam"], consts["mu"], consts["typ"]
    dt, dx = consts["dt"], consts["dx"]
    lo, hi, dims = consts["lo"], consts["hi"], consts["dims"]
    g = consts["g"]
    invdx = consts.get("invdx", 1.0 / dx)
    n = x.shape[0]
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
    # P2G
    vp = v + g * dt
    tau = stress(F, lam, mu, typ)
    sc = -dt * 4.0 * invdx * invdx * V0
    A = tau * sc[:, None, None] + C * m[:, None, None]
    contrib = (m[:, None, None] * vp[:, None, :] + torch.einsum("nrc,nkc->nkr", A, dist)) * wt_in[..., None]
    nodes = nx * ny * nz
    P = torch.zeros(nodes, 3, dtype=x.dtype).index_add(0, flat.reshape(-1), contrib.reshape(-1, 3))
    M = torch.zeros(nodes, dtype=x.dtype).index_add(0, flat.reshape(-1), (wt_in * m[:, None]).reshape(-1))
    # grid update
    heavy = M >= 1e-4
    U = torch.where(heavy[:, None], P / torch.where(heavy, M, torch.ones_like(M))[:, None], P)
    idx = torch.arange(nodes)
    ii, jj, kk = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    band = torch.stack([(ii < 2) | (ii > nx - 3), (jj < 2) | (jj > ny - 3), (kk < 2) | (kk > nz - 3)], 1)
    U = torch.where(heavy[:, None] & band, torch.zeros_like(U), U)
    # G2P
    Ui = U[flat] * inside[..., None]
    vnew = (Ui * wt[..., None]).sum(1)
    B = torch.einsum("nk,nkr,nkc->nrc", wt, dist, Ui)
    Cn = 4.0 * invdx * B
    xnew = torch.minimum(torch.maximum(x + dt * vnew, lo), hi)
    eye = torch.eye(3, dtype=x.dtype).expand(n, 3, 3)
    Fnew = (eye + dt * Cn) @ F
    return xnew, vnew, Fnew, Cn


def rollout(x0, v0, F0, C0, consts, steps):
    state = (x0, v0, F0, C0)
    for _ in range(steps):
        state = substep(state, consts)
    return state


def consts_to(consts, dtype):
    """The same constants in another float type (an FP32 evaluation of this
    restatement measures how much of a gradient difference FP32 explains)."""
    out = dict(consts)
    for k in ("m", "V0", "lam", "mu", "g", "lo", "hi"):
        out[k] = consts[k].to(dtype)
    return out


def consts_from_scene(scene, V0, lam=None, mu=None):
    """Constant inputs of a scene (masses, volumes, materials, grid) in FP64."""
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    dims = [int(x) for x in scene["dims"]] if "dims" in scene else None
    if dims is None:
        ext = scene["hi"] - scene["lo"]
        dims = [int(np.float32(e) / np.float32(scene["dx"]) + np.float32(1.0)) for e in ext]
    # the solver interfaces take dt, dx and gravity as FP32 (include/solver/mpm_solver.h:63-190)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    return dict(m=t(scene["xm"][:, 3]), V0=t(V0), lam=t(scene["params"][:, 0]) if lam is None else lam,
                mu=t(scene["params"][:, 1]) if mu is None else mu, typ=torch.as_tensor(scene["type"]),
                g=t(f32(scene["gravity"])), dt=float(f32(scene["dt"])), dx=float(f32(scene["dx"])),
                # the solvers hold 1/dx as an FP32 value (src/solver/gpu_mls_mpm_solver.cu:40-52)
                invdx=float(np.float32(1.0) / np.float32(scene["dx"])),
                lo=t(scene["lo"]), hi=t(scene["hi"]), dims=torch.tensor(dims))
