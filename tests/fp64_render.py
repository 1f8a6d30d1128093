"""FP64 torch re-render of fixed composited hit lists, differentiable in the
PLY-raw parameters: the reference for the backward kernel.

Pixel of a ray with composited hits h_1..h_n (front to back):
    L = sum_i T_i alpha_i c_i,  T_1 = 1, T_{i+1} = T_i (1 - alpha_i)
with the activations of GS_ply.ts:225-235 (scale = exp, quaternion
normalised, opacity = sigmoid) evaluated in double from the raw values,
alpha = computeResponse * opacity with computeResponse written in the
reference's covariance form (GS_shader.ts:111-136: invCov = M M^T of the
rotation-over-scale matrix, the closest point s* = (mu-o).invCov.d /
d.invCov.d, response exp(-(p-mu).invCov.(p-mu))), and c = max(0, 0.5 + SH
(GS_shader.ts:138-190) at the direction normalize(mu - o)).  This is not
the kernel's formulation (whitened cross product), so agreement pins both.
"""
import numpy as np
import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435]


def sh_basis(d):
    """(m, 16) real SH basis of unit directions d (m, 3), GS_shader.ts:142-175."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    b = [torch.full_like(x, C0), -C1 * y, C1 * z, -C1 * x,
         C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z, C2[4] * (xx - yy),
         C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
         C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
         C3[6] * x * (xx - 3 * yy)]
    return torch.stack(b, dim=1)


def flatten_checkpoint(comp_ids, n_comp):
    """Ragged (ray, slot) -> (ray index, gid, position) arrays of the composited hits."""
    n_comp = np.asarray(n_comp, np.int64)
    rays = np.repeat(np.arange(len(n_comp)), n_comp)
    pos = np.concatenate([np.arange(c) for c in n_comp]) if len(n_comp) else np.zeros(0, np.int64)
    gids = np.asarray(comp_ids)[rays, pos].astype(np.int64)
    return rays, gids, pos


def render(raw, o, dirs, rays, gids, pos, n_rays, sh_degree=3):
    """raw (n, 59) float64 tensor; o (3,) and dirs (n_rays, 3) float64 tensors;
    rays/gids/pos from flatten_checkpoint.  Returns (n_rays, 3) float64."""
    dev = raw.device
    rays_t = torch.as_tensor(rays, device=dev)
    gids_t = torch.as_tensor(gids, device=dev)
    pos_t = torch.as_tensor(pos, device=dev)
    p = raw[gids_t]
    mu = p[:, 0:3]
    sc = torch.exp(p[:, 52:55])
    q = p[:, 55:59] / torch.linalg.norm(p[:, 55:59], dim=1, keepdim=True)
    op = torch.sigmoid(p[:, 51])
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    # shader mat3 R (column-major R[c][r]) of the quaternion; M[c][r] = R[c][r] / s[r]
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)], 1),
        torch.stack([2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)], 1),
        torch.stack([2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    M = R / sc[:, None, :]
    inv_cov = torch.einsum("mck,mrk->mcr", M, M)          # A[c][r] = sum_k M[r][k] M[c][k]
    d = dirs[rays_t]
    tmp = torch.einsum("mcr,mc->mr", inv_cov, d)
    mo = mu - o[None, :]
    s = (mo * tmp).sum(1) / (d * tmp).sum(1)
    pm = o[None, :] + s[:, None] * d - mu
    qv = torch.einsum("mcr,mc->mr", inv_cov, pm)
    alpha = torch.exp(-(pm * qv).sum(1)) * op
    # colour
    w = mu - o[None, :]
    dirn = w / torch.linalg.norm(w, dim=1, keepdim=True)
    nb = {0: 1, 1: 4, 2: 9, 3: 16}[sh_degree]
    basis = sh_basis(dirn)[:, :nb]
    shc = torch.cat([p[:, 3:6][:, None, :], p[:, 6:51].reshape(-1, 3, 15).transpose(1, 2)], dim=1)[:, :nb]
    col = torch.clamp(0.5 + (basis[:, :, None] * shc).sum(1), min=0.0)
    # transmittance per ray: padded (n_rays, max_n) table
    maxn = int(pos.max()) + 1 if len(pos) else 1
    A = torch.zeros((n_rays, maxn), dtype=raw.dtype, device=dev)
    A = A.index_put((rays_t, pos_t), alpha)
    one_m = 1.0 - A
    T = torch.cat([torch.ones((n_rays, 1), dtype=raw.dtype, device=dev), torch.cumprod(one_m, dim=1)[:, :-1]], 1)
    wgt = T[rays_t, pos_t] * alpha
    out = torch.zeros((n_rays, 3), dtype=raw.dtype, device=dev)
    return out.index_add(0, rays_t, wgt[:, None] * col)
