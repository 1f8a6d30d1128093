"""GPU tests of the hand-written backward kernel (gt_backward).

Reference: tests/fp64_render.py, an FP64 torch re-render of the same
composited hit lists in the reference's covariance form, differentiated by
torch autograd and by central finite differences (step 1e-4).  The hit
lists are held fixed (the forward's list changes have no derivative).
Bar: per parameter group relative L2 error <= 1e-3 and per element
|g - ref| <= 1e-3 (|ref| + 1e-3 max|ref_group|).
"""
import numpy as np
import pytest
import torch

from helpers import uniforms
import fp64_render as fr
import gtrace as gt
import gtrace.autograd as gta
import oracle as orc

pytestmark = pytest.mark.gpu

GROUPS = {"mean": slice(0, 3), "f_dc": slice(3, 6), "f_rest": slice(6, 51), "logit": slice(51, 52),
          "log_scale": slice(52, 55), "quat": slice(55, 59)}
REL_TOL = 1e-3


def _print(*a):
    print(*a, flush=True)


def _case(n=3000, W=48, H=40, seed=3, k=8, sh=3, alpha_min=0.05, pos=(0.3, -0.2, 5.0), **gen):
    raw = gt.generate_raw(n=n, seed=seed, config="c1", rest_std=0.2, **gen)
    s = gt.Scene(raw, alpha_min=alpha_min)
    u = uniforms(list(pos), [0, 0, 0], W, H, hfov_deg=50, alpha_min=alpha_min, k=k, T_min=1e-3, sh_degree_max=sh)
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((H, W, 4)).astype(np.float32)
    G[..., 3] = 0.0
    return raw, s, u, G


def _gpu_grads(s, n, u, W, H, G, mode="checkpoint", cap=256, budget=0):
    dev = torch.device("cuda", 0)
    g = torch.from_numpy(G).to(dev)
    grads = torch.zeros((n, 59), dtype=torch.float32, device=dev)
    ovf = torch.zeros(1, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    if mode == "checkpoint":
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
        cids = torch.empty((H * W, cap), dtype=torch.int32, device=dev)
        nc = torch.empty(H * W, dtype=torch.int32, device=dev)
        s.render_device(u, W, H, rgba.data_ptr(), st, comp_ids_ptr=cids.data_ptr(), n_comp_ptr=nc.data_ptr(),
                        comp_cap=cap)
        gt.backward(s, u, W, H, g.data_ptr(), grads.data_ptr(), cids.data_ptr(), nc.data_ptr(), cap, ovf.data_ptr(),
                    0, st)
    else:
        gt.backward(s, u, W, H, g.data_ptr(), grads.data_ptr(), None, None, cap, ovf.data_ptr(), budget, st)
    torch.cuda.synchronize()
    return grads.cpu().numpy().astype(np.float64), int(ovf.cpu().item())


def _reference(raw, s, u, W, H, G, sh):
    """(fp64 pixels, autograd grads, checkpoint, closure f(raw64)->loss)."""
    ck = s.render(u, W, H, checkpoint_cap=256)
    R = W * H
    rays, gids, pos = fr.flatten_checkpoint(ck["comp_ids"], ck["n_comp"])
    p = orc.params_from_uniforms(u, W, H)
    dirs = torch.from_numpy(np.stack([orc.ray_dir(p, r % W, r // W) for r in range(R)]).astype(np.float64))
    o = torch.tensor([u.to_world[12], u.to_world[13], u.to_world[14]], dtype=torch.float64)
    Gt = torch.from_numpy(G.reshape(R, 4)[:, :3].astype(np.float64))

    def loss_of(raw64):
        out = fr.render(raw64, o, dirs, rays, gids, pos, R, sh_degree=sh)
        return (out * Gt).sum(), out

    raw_t = torch.from_numpy(raw.astype(np.float64)).requires_grad_(True)
    loss, out = loss_of(raw_t)
    loss.backward()
    return out.detach().numpy(), raw_t.grad.numpy(), ck, loss_of


def _check_grads(g, ref, label):
    for name, sl in GROUPS.items():
        a, b = g[:, sl], ref[:, sl]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.abs(a).max() == 0, f"{label} {name}: nonzero gradient where the reference is zero"
            continue
        rel = np.linalg.norm(a - b) / nb
        scale = np.abs(b).max()
        el = np.abs(a - b) / (np.abs(b) + 1e-3 * scale)
        _print(f"  {label} {name}: rel L2 {rel:.2e}, worst element ratio {el.max():.2e}")
        assert rel <= REL_TOL, f"{label} {name}: relative L2 error {rel:.3e}"
        assert el.max() <= REL_TOL, f"{label} {name}: element error {el.max():.3e}"


def test_checkpoint_matches_oracle_hit_lists():
    """The composited ids of the checkpoint are the oracle's accepted hits, in order."""
    _print("case backward checkpoint vs oracle 48x40 k8")
    raw, s, u, _ = _case()
    W, H = 48, 40
    ck = s.render(u, W, H, checkpoint_cap=256)
    o = orc.OracleScene(raw, 0.05)
    r = orc.render(o, orc.params_from_uniforms(u, W, H), max_rounds=64)
    assert r["stats"]["overflow"] == 0
    bits = (r["accept"][:, :, None] >> np.arange(u.k, dtype=np.uint32)) & 1
    for ray in range(W * H):
        want = r["hit_ids"][ray][bits[ray].astype(bool)]
        n = int(ck["n_comp"][ray])
        assert n == len(want), f"ray {ray}: {n} composited, oracle {len(want)}"
        np.testing.assert_array_equal(ck["comp_ids"][ray, :n], want, err_msg=f"ray {ray}")
    assert int(ck["n_comp"].sum()) == r["stats"]["h_comp"] > 500


@pytest.mark.parametrize("sh", [3, 1])
def test_backward_vs_fp64_autograd(sh):
    _print(f"case backward vs fp64 autograd sh{sh}")
    raw, s, u, G = _case(sh=sh)
    W, H = 48, 40
    rgb64, ref, ck, _ = _reference(raw, s, u, W, H, G, sh)
    rgba = s.render(u, W, H)["rgba"].reshape(-1, 4)
    err = np.abs(rgba[:, :3] - rgb64).max()
    _print(f"  fp64 re-render vs kernel pixels: {err:.2e}")
    assert err <= 1e-4
    g, ovf = _gpu_grads(s, raw.shape[0], u, W, H, G)
    assert ovf == 0
    _check_grads(g, ref, f"sh{sh}")
    if sh == 1:
        used = [6 + c * 15 + j for c in range(3) for j in range(3)]
        unused = [c for c in range(6, 51) if c not in used]
        assert np.all(g[:, unused] == 0) and np.any(g[:, used] != 0)


def test_backward_vs_central_differences():
    """512 parameters of Gaussians on the hit lists, central FD of the FP64 re-render, step 1e-4."""
    _print("case backward vs central FD 512 params")
    raw, s, u, G = _case()
    W, H = 48, 40
    _, ref, ck, loss_of = _reference(raw, s, u, W, H, G, 3)
    g, _ = _gpu_grads(s, raw.shape[0], u, W, H, G)
    hit = np.unique(np.concatenate([ck["comp_ids"][r, :ck["n_comp"][r]] for r in range(W * H)]))
    rng = np.random.default_rng(7)
    picks = []
    per = {"mean": 96, "f_dc": 64, "f_rest": 96, "logit": 64, "log_scale": 96, "quat": 96}
    for name, cnt in per.items():
        sl = GROUPS[name]
        cols = np.arange(sl.start, sl.stop)
        picks += [(int(rng.choice(hit)), int(rng.choice(cols))) for _ in range(cnt)]
    assert len(picks) == 512
    raw64 = torch.from_numpy(raw.astype(np.float64))
    h = 1e-4
    fd = np.zeros(len(picks))
    with torch.no_grad():
        for i, (gid, col) in enumerate(picks):
            rp = raw64.clone()
            rp[gid, col] += h
            rm = raw64.clone()
            rm[gid, col] -= h
            fd[i] = (loss_of(rp)[0].item() - loss_of(rm)[0].item()) / (2 * h)
    gv = np.array([g[a, b] for a, b in picks])
    av = np.array([ref[a, b] for a, b in picks])
    scale = np.abs(fd).max()
    tol = 1e-3 * (np.abs(fd) + 1e-2 * scale)
    bad = np.nonzero(np.abs(gv - fd) > tol)[0]
    _print(f"  FD max |kernel-fd| {np.abs(gv - fd).max():.2e}, |autograd-fd| {np.abs(av - fd).max():.2e}, "
           f"max|fd| {scale:.2e}")
    assert bad.size == 0, f"{bad.size} of 512 differ, first {[picks[i] for i in bad[:4]]} {gv[bad[:4]]} {fd[bad[:4]]}"
    assert np.all(np.abs(av - fd) <= tol)


def test_backward_shared_gaussians_ragged_tiles():
    """Gaussian 0 lies in front of the camera and covers the view, so every
    ray composites it first; Gaussian 1 covers about half of it.  The lanes
    of a wave then hold the same Gaussian at the same step, which is the case
    where the kernel merges lanes before its LDS atomics.  45x37 leaves
    ragged 8x8 tiles at the right and bottom edges."""
    _print("case backward shared Gaussians, ragged tiles 45x37")
    raw = gt.generate_raw(n=2000, seed=5, config="c1", rest_std=0.2)
    raw[0, 0:3] = [0.25, -0.15, 2.8]
    raw[0, 52:55] = np.log(0.8)
    raw[0, 51] = 0.0
    raw[0, 55:59] = [1.0, 0.0, 0.0, 0.0]
    raw[1, 0:3] = [0.9, -0.3, 2.6]
    raw[1, 52:55] = np.log([0.5, 0.8, 0.4])
    raw[1, 51] = 0.5
    raw[1, 55:59] = [0.9, 0.2, -0.3, 0.1]
    W, H = 45, 37
    s = gt.Scene(raw, alpha_min=0.05)
    u = uniforms([0.3, -0.2, 5.0], [0, 0, 0], W, H, hfov_deg=50, alpha_min=0.05, k=8, T_min=1e-3, sh_degree_max=3)
    G = np.random.default_rng(5).standard_normal((H, W, 4)).astype(np.float32)
    G[..., 3] = 0.0
    _, ref, ck, _ = _reference(raw, s, u, W, H, G, 3)
    first = ck["comp_ids"][:, 0][ck["n_comp"] > 0]
    both = np.mean([1 in ck["comp_ids"][r, :ck["n_comp"][r]] for r in range(W * H)])
    _print(f"  rays with Gaussian 0 first: {np.mean(first == 0):.3f}, with Gaussian 1: {both:.3f}")
    assert np.all(first == 0) and 0.2 < both < 0.8
    g, ovf = _gpu_grads(s, raw.shape[0], u, W, H, G)
    assert ovf == 0
    _check_grads(g, ref, "shared")


def test_backward_block_table_overflow():
    """20k pixel-sized Gaussians: each 16x16 block of rays composites about a
    thousand distinct Gaussians, more than the block's 512-entry LDS table
    holds, so many accumulators take the direct HBM atomic path."""
    _print("case backward LDS table overflow 64x48")
    raw = gt.generate_raw(n=20000, seed=11, config="c1", rest_std=0.2, log_scale_mean=-3.2, log_scale_std=0.3)
    W, H = 64, 48
    s = gt.Scene(raw, alpha_min=0.05)
    u = uniforms([0.3, -0.2, 5.0], [0, 0, 0], W, H, hfov_deg=50, alpha_min=0.05, k=8, T_min=1e-3, sh_degree_max=3)
    G = np.random.default_rng(11).standard_normal((H, W, 4)).astype(np.float32)
    G[..., 3] = 0.0
    _, ref, ck, _ = _reference(raw, s, u, W, H, G, 3)
    ids = ck["comp_ids"].reshape(H, W, -1)
    nc = ck["n_comp"].reshape(H, W)
    distinct = [len({int(i) for y in range(by, by + 16) for x in range(bx, bx + 16) for i in ids[y, x, :nc[y, x]]})
                for by in range(0, H, 16) for bx in range(0, W, 16)]
    _print(f"  distinct Gaussians per 16x16 block: min {min(distinct)}, max {max(distinct)}")
    assert min(distinct) > 1.5 * 512
    g, ovf = _gpu_grads(s, raw.shape[0], u, W, H, G)
    assert ovf == 0
    _check_grads(g, ref, "overflow")


def test_recompute_equals_checkpoint():
    """Recompute mode over several bands (small scratch budget) == checkpoint mode up to atomic order."""
    _print("case backward recompute vs checkpoint")
    raw, s, u, G = _case(W=72, H=56)
    W, H = 72, 56
    n = raw.shape[0]
    a, _ = _gpu_grads(s, n, u, W, H, G, "checkpoint")
    per_row = W * 8 * 257 * 4
    for budget, cap in ((2 * per_row, 256), (0, 256), (W * 8 * 2 * 4, 2)):
        # the last case starts with room for 2 ids per ray: the scratch must grow, never overflow
        b, ovf = _gpu_grads(s, n, u, W, H, G, "recompute", cap=cap, budget=budget)
        assert ovf == 0
        np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-6 * np.abs(a).max())


def test_checkpoint_overflow_is_reported():
    _print("case backward overflow")
    raw, s, u, G = _case()
    W, H = 48, 40
    ck = s.render(u, W, H, checkpoint_cap=256)
    _, ovf = _gpu_grads(s, raw.shape[0], u, W, H, G, cap=2)
    assert ovf == int((ck["n_comp"] > 2).sum()) > 0


def test_autograd_function(monkeypatch):
    _print("case torch.autograd wrapper")
    monkeypatch.setattr(gta, "CHECKPOINT_CAP", 2)  # forces the re-render with a larger checkpoint
    raw, s, u, G = _case()
    W, H = 48, 40
    dev = torch.device("cuda", 0)
    want, _ = _gpu_grads(s, raw.shape[0], u, W, H, G)
    rgba_host = s.render(u, W, H)["rgba"]
    for mode in ("checkpoint", "recompute"):
        x = torch.from_numpy(raw).to(dev).requires_grad_(True)
        rgba = gta.render(x, u, W, H, 0.05, mode=mode)
        np.testing.assert_array_equal(rgba.detach().cpu().numpy().reshape(rgba_host.shape), rgba_host)
        (rgba * torch.from_numpy(G).to(dev)).sum().backward()
        np.testing.assert_allclose(x.grad.cpu().numpy(), want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_backward_rejects_bad_arguments():
    raw, s, u, G = _case(n=200)
    dev = torch.device("cuda", 0)
    g = torch.zeros((40, 48, 4), device=dev)
    grads = torch.zeros((200, 59), device=dev)
    cids = torch.zeros((48 * 40, 4), dtype=torch.int32, device=dev)
    with pytest.raises(gt.GTError):
        gt.backward(s, u, 48, 40, g.data_ptr(), grads.data_ptr(), cids.data_ptr(), None, 4)
    u2 = uniforms([0, 0, 5], [0, 0, 0], 48, 40, debug_mode=1)
    with pytest.raises(gt.GTError):
        gt.backward(s, u2, 48, 40, g.data_ptr(), grads.data_ptr())


def test_count_only_checkpoint():
    """n_comp without comp_ids counts the composited hits (to size a checkpoint)."""
    _print("case count-only checkpoint")
    raw, s, u, _ = _case()
    full = s.render(u, 48, 40, checkpoint_cap=256)
    cnt = s.render(u, 48, 40, checkpoint_cap=-1)
    np.testing.assert_array_equal(cnt["n_comp"], full["n_comp"])
    np.testing.assert_array_equal(cnt["rgba"], full["rgba"])


C4_CAMERA = ([0.0, 0.5, 4.0], [0.0, 0.0, 0.0])


def test_c4_scale_fd_512_and_modes():
    """200k single-cluster Gaussians, one 800x800 camera, alpha_min 0.01, k 16,
    T_min 0.001; loss = sum (render - target)^2 with the target rendered from
    means + N(0, 0.01).  Kernel gradients vs FP64 central differences (step
    1e-4) of the FP64 re-render of the frozen hit lists on 512 seeded
    parameters (seed 2), and checkpoint vs recompute mode."""
    _print("case C4 200k 800x800 k16: FD 512 + checkpoint vs recompute")
    W = H = 800
    raw = gt.generate_raw(seed=0, config="c4")
    n = raw.shape[0]
    s = gt.Scene(raw, alpha_min=0.01)
    u = uniforms(C4_CAMERA[0], C4_CAMERA[1], W, H, hfov_deg=50, alpha_min=0.01, k=16, T_min=1e-3)
    rng = np.random.default_rng(3)
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    target = gt.Scene(raw_t, alpha_min=0.01).render(u, W, H)["rgba"].reshape(H * W, 4)
    _print(f"  C4: target rendered")
    cnt = s.render(u, W, H, checkpoint_cap=-1)["n_comp"]
    cap = 1 << int(max(1, cnt.max()) - 1).bit_length()
    ck = s.render(u, W, H, checkpoint_cap=cap)
    R = W * H
    rays, gids, pos = fr.flatten_checkpoint(ck["comp_ids"], ck["n_comp"])
    p = orc.params_from_uniforms(u, W, H)
    dirs_all = np.stack([orc.ray_dir(p, r % W, r // W) for r in range(R)]).astype(np.float64)
    o = torch.tensor([u.to_world[12], u.to_world[13], u.to_world[14]], dtype=torch.float64)
    raw64 = torch.from_numpy(raw.astype(np.float64))
    tgt64 = torch.from_numpy(target[:, :3].astype(np.float64))
    # FP64 pixels of the frozen hit lists, in chunks of rays; the loss
    # gradient dL/dRGBA = 2 (render64 - target) is handed to the kernel
    out64 = np.zeros((R, 3))
    bounds = list(range(0, R, 65536)) + [R]
    starts = np.searchsorted(rays, bounds)
    with torch.no_grad():
        for c in range(len(bounds) - 1):
            a, b = starts[c], starts[c + 1]
            r0 = bounds[c]
            nr = bounds[c + 1] - r0
            out64[r0:r0 + nr] = fr.render(raw64, o, torch.from_numpy(dirs_all[r0:r0 + nr]), rays[a:b] - r0,
                                          gids[a:b], pos[a:b], nr).numpy()
    assert np.abs(out64 - ck["rgba"].reshape(R, 4)[:, :3]).max() <= 1e-4
    G = np.zeros((R, 4), np.float32)
    G[:, :3] = 2.0 * (out64 - target[:, :3].astype(np.float64))
    G = G.reshape(H, W, 4)
    g_ck, ovf = _gpu_grads(s, n, u, W, H, G, "checkpoint", cap=cap)
    assert ovf == 0
    g_rc, ovf = _gpu_grads(s, n, u, W, H, G, "recompute")
    assert ovf == 0
    np.testing.assert_allclose(g_rc, g_ck, rtol=1e-5, atol=1e-6 * np.abs(g_ck).max())
    # FD on the rays that composite each picked Gaussian
    order = np.argsort(gids, kind="stable")
    gs, rs = gids[order], rays[order]
    hit = np.unique(gids)
    pick_rng = np.random.default_rng(2)
    per = {"mean": 96, "f_dc": 64, "f_rest": 96, "logit": 64, "log_scale": 96, "quat": 96}
    picks = []
    for name, cnt_ in per.items():
        sl = GROUPS[name]
        picks += [(int(pick_rng.choice(hit)), int(pick_rng.integers(sl.start, sl.stop))) for _ in range(cnt_)]
    h = 1e-4
    fd = np.zeros(len(picks))
    with torch.no_grad():
        for i, (gid, col) in enumerate(picks):
            lo, hi = np.searchsorted(gs, gid), np.searchsorted(gs, gid, side="right")
            sub = np.unique(rs[lo:hi])
            srays, sgids, spos = fr.flatten_checkpoint(ck["comp_ids"][sub], ck["n_comp"][sub])
            dirs = torch.from_numpy(dirs_all[sub])
            tsub = tgt64[torch.from_numpy(sub)]

            def loss():
                out = fr.render(raw64, o, dirs, srays, sgids, spos, len(sub))
                return float(((out - tsub) ** 2).sum())

            raw64[gid, col] += h
            lp = loss()
            raw64[gid, col] -= 2 * h
            lm = loss()
            raw64[gid, col] += h
            fd[i] = (lp - lm) / (2 * h)
    gv = np.array([g_ck[a, b] for a, b in picks])
    scale = np.abs(fd).max()
    err = np.abs(gv - fd)
    tol = 1e-3 * (np.abs(fd) + 1e-2 * scale)
    _print(f"  C4 FD: max |kernel - fd| {err.max():.3e}, max |fd| {scale:.3e}, worst ratio {(err / tol).max():.3f}")
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, f"{bad.size} of 512 differ, first {[picks[i] for i in bad[:4]]} {gv[bad[:4]]} {fd[bad[:4]]}"
