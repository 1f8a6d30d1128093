"""The FP64 torch re-render used as the backward reference, pinned on the CPU
against the oracle's FP64 re-evaluation of the oracle's own hit lists (a
different formulation of the same pixel), and its autograd against central
differences."""
import numpy as np
import torch

from helpers import uniforms
import fp64_render as fr
import gtrace as gt
import oracle as orc


def _oracle_case(sh=3):
    raw = gt.generate_raw(n=1500, seed=5, config="c1", rest_std=0.2)
    W, H = 24, 20
    u = uniforms([0.2, 0.1, 5.0], [0, 0, 0], W, H, hfov_deg=50, alpha_min=0.05, k=8, T_min=1e-3, sh_degree_max=sh)
    o = orc.OracleScene(raw, 0.05)
    p = orc.params_from_uniforms(u, W, H)
    r = orc.render(o, p, max_rounds=64)
    bits = (r["accept"][:, :, None] >> np.arange(8, dtype=np.uint32)) & 1
    R = W * H
    lists = [r["hit_ids"][i][bits[i].astype(bool)] for i in range(R)]
    n = np.array([len(x) for x in lists])
    cids = np.zeros((R, max(1, n.max())), np.uint32)
    for i, x in enumerate(lists):
        cids[i, :len(x)] = x
    rays, gids, pos = fr.flatten_checkpoint(cids, n)
    dirs = torch.from_numpy(np.stack([orc.ray_dir(p, i % W, i // W) for i in range(R)]).astype(np.float64))
    ov = torch.tensor([u.to_world[12], u.to_world[13], u.to_world[14]], dtype=torch.float64)
    return raw, o, p, r, rays, gids, pos, dirs, ov, R


def test_fp64_render_matches_oracle_reeval():
    for sh in (3, 0):
        raw, o, p, r, rays, gids, pos, dirs, ov, R = _oracle_case(sh)
        rgb64, _ = orc.reeval_f64(o, p, r["n_rounds"], r["hit_ids"], r["accept"])
        out = fr.render(torch.from_numpy(raw.astype(np.float64)), ov, dirs, rays, gids, pos, R, sh_degree=sh)
        assert len(rays) > 200
        # the oracle evaluates FP32-packed activations with double ray directions
        assert np.abs(out.numpy() - rgb64).max() < 2e-5


def test_fp64_render_autograd_vs_differences():
    raw, o, p, r, rays, gids, pos, dirs, ov, R = _oracle_case()
    rng = np.random.default_rng(1)
    w = torch.from_numpy(rng.standard_normal((R, 3)))
    x = torch.from_numpy(raw.astype(np.float64)).requires_grad_(True)
    (fr.render(x, ov, dirs, rays, gids, pos, R) * w).sum().backward()
    hit = np.unique(gids)
    base = torch.from_numpy(raw.astype(np.float64))
    with torch.no_grad():
        for _ in range(24):
            gid, col = int(rng.choice(hit)), int(rng.integers(0, 59))
            a, b = base.clone(), base.clone()
            a[gid, col] += 1e-5
            b[gid, col] -= 1e-5
            fd = ((fr.render(a, ov, dirs, rays, gids, pos, R) * w).sum() -
                  (fr.render(b, ov, dirs, rays, gids, pos, R) * w).sum()).item() / 2e-5
            assert abs(fd - x.grad[gid, col].item()) <= 1e-6 * (1 + abs(fd)), (gid, col)
