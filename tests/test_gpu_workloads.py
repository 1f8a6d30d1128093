"""GPU parity at the benchmark workloads themselves, through the C-ABI.

- 1M Gaussians, 1920x1080, k 16 (the headline frame): the whole frame's
  hit lists, accept masks, round counts and T are dumped on the device; a
  seeded 1/64 ray sample is checked against the oracle bit for bit, with
  pixels within 1e-4 of the FP64 re-evaluation, and the full frame against
  its invariants and its own counters.
- 64 cameras of uneven resolution on the same scene in one graph-replayed
  launch, for alpha_min 0.01 and 0.1 and k 4, 8, 16, 32: every camera
  equals gt_render of that camera bit for bit (and again after autotuning),
  and 256 seeded rays per camera equal the oracle's hit lists.
- 3M Gaussians, 16 views at 1280x720, gradients accumulated across views:
  checkpoint mode equals recompute mode within FP32-atomic tolerance, and
  512 seeded parameters match the FP64 gradient (autograd) of the FP64
  re-render of the frozen hit lists within 1e-3 relative, which central
  differences (steps 1e-4 and 5e-5, Richardson-extrapolated) confirm.
The reference's round loop is src/GS_shader.ts:192-273; the batch replaces
one traceRays dispatch per camera (webrtx/src/pass_encoder.ts:35-65).
"""
import os
import sys

import numpy as np
import pytest
import torch

import fp64_render as fr
import gtrace as gt
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workloads' seeded cameras)

pytestmark = pytest.mark.gpu

PIXEL_TOL = 1e-4
DEV = torch.device("cuda", 0)
SENT = -1  # GT_SENTINEL_ID as int32


def _print(*a):
    print(*a, flush=True)


_CACHE = {}


def _c2_raw():
    if "raw" not in _CACHE:
        _CACHE["raw"] = gt.generate_raw(seed=0, config="c2")
    return _CACHE["raw"]


def _c2_pair(alpha):
    """(library scene, oracle scene with its BVH) of the 1M-Gaussian scene, built once per alpha_min."""
    key = ("pair", alpha)
    if key not in _CACHE:
        raw = _c2_raw()
        s = gt.Scene(raw, alpha_min=alpha)
        o = orc.OracleScene(raw, alpha)
        o.prebuild_bvh()
        _CACHE[key] = (s, o)
    return _CACHE[key]


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _dump_render(s, u, W, H, max_rounds, stats=False):
    """One render with every per-ray output on the device."""
    R = W * H
    out = dict(rgba=torch.empty((H, W, 4), dtype=torch.float32, device=DEV),
               n_rounds=torch.empty(R, dtype=torch.int32, device=DEV),
               T=torch.empty(R, dtype=torch.float32, device=DEV))
    if max_rounds:
        out["hit_ids"] = torch.empty((R, max_rounds, u.k), dtype=torch.int32, device=DEV)
        out["accept"] = torch.empty((R, max_rounds), dtype=torch.int32, device=DEV)
    if stats:
        out["stats"] = torch.zeros(gt.GT_NSTATS, dtype=torch.int64, device=DEV)
    s.render_device(u, W, H, out["rgba"].data_ptr(), _stream(),
                    stats_ptr=out["stats"].data_ptr() if stats else None,
                    hit_ids_ptr=out["hit_ids"].data_ptr() if max_rounds else None,
                    accept_ptr=out["accept"].data_ptr() if max_rounds else None,
                    n_rounds_ptr=out["n_rounds"].data_ptr(), T_ptr=out["T"].data_ptr(), max_rounds=max_rounds)
    return out


def _check_sample(g, o, u, W, H, idx, label):
    """The device outputs g at rays idx against the oracle on the same rays."""
    MR = g["hit_ids"].shape[1]
    it = torch.from_numpy(idx.astype(np.int64)).to(DEV)
    hits = g["hit_ids"][it].cpu().numpy().view(np.uint32)
    acc = g["accept"][it].cpu().numpy().view(np.uint32)
    rounds = g["n_rounds"][it].cpu().numpy().view(np.uint32)
    T = g["T"][it].cpu().numpy()
    rgba = g["rgba"].reshape(-1, 4)[it].cpu().numpy()
    p = orc.params_from_uniforms(u, W, H)
    r = orc.render(o, p, max_rounds=MR, ray_list=idx.astype(np.uint32))
    assert r["stats"]["overflow"] == 0, f"{label}: oracle needs more rounds than {MR}"
    np.testing.assert_array_equal(rounds, r["n_rounds"], err_msg=f"{label}: rounds")
    bad = np.nonzero((hits != r["hit_ids"]).any(axis=(1, 2)))[0]
    assert bad.size == 0, f"{label}: hit lists differ on {bad.size} rays, first rays {idx[bad[:5]]}"
    np.testing.assert_array_equal(acc, r["accept"], err_msg=f"{label}: accept masks")
    np.testing.assert_array_equal(T.view(np.uint32), r["T"].view(np.uint32), err_msg=f"{label}: T bits")
    rgb64, _ = orc.reeval_f64(o, p, rounds, hits, acc, ray_list=idx.astype(np.uint32))
    err = float(np.abs(rgba[:, :3].astype(np.float64) - rgb64).max()) if len(idx) else 0.0
    assert err <= PIXEL_TOL, f"{label}: pixel error {err}"
    assert np.all(rgba[:, 3] == 1.0)
    return r, err


def _popcount_sum(acc):
    a = acc.to(torch.int64) & 0xFFFFFFFF
    return int(sum(((a >> b) & 1).sum().item() for b in range(32)))


def test_c2_full_workload():
    """1,000,000 Gaussians (64-cluster mixture, scene seed 0), one 1920x1080
    camera at the seeded position inside the volume looking at the origin,
    hFov 50, alpha_min 0.01, k 16, T_min 0.001: the benchmark's frame."""
    _print("case C2 1M 1920x1080 k16: full-frame dump, 1/64 sample vs oracle")
    W, H, K = 1920, 1080, 16
    R = W * H
    s, o = _c2_pair(0.01)
    pos, tgt = bench.camera_c2()
    u = gt.make_uniforms(gt.camera_look_at(pos, tgt), W, H, hfov_deg=50.0, t_min=1e-3, t_max=1e5, T_min=1e-3,
                         alpha_min=0.01, k=K, sh_degree_max=3)
    first = _dump_render(s, u, W, H, 0, stats=True)
    torch.cuda.synchronize()
    MR = int(first["n_rounds"].max().item())
    _print(f"  C2: at most {MR} rounds per ray")
    g = _dump_render(s, u, W, H, MR, stats=True)
    torch.cuda.synchronize()
    st = dict(zip(gt.STAT_NAMES, g["stats"].cpu().tolist()))
    # the frame with dumps equals the frame without, bit for bit
    assert torch.equal(g["rgba"], first["rgba"]) and torch.equal(g["n_rounds"], first["n_rounds"])
    assert torch.equal(g["T"].view(torch.int32), first["T"].view(torch.int32))
    # full-frame invariants and the kernel's own counters against the dumps
    assert st["rays"] == R and st["overflow"] == 0
    assert st["rounds"] == int(g["n_rounds"].to(torch.int64).sum().item())
    assert int(g["n_rounds"].min().item()) >= 1
    assert st["h_eval"] == int((g["hit_ids"] != SENT).sum().item())
    assert st["h_comp"] == _popcount_sum(g["accept"])
    assert bool((g["rgba"][..., 3] == 1.0).all()) and bool((g["rgba"][..., :3] >= 0).all())
    assert bool((g["T"] > 0).all()) and bool((g["T"] <= 1).all())
    # a finished ray either fell to T_min or ran out of hits (its last round has a sentinel)
    last = g["hit_ids"][torch.arange(R, device=DEV), (g["n_rounds"] - 1).long()]
    ran_out = (last == SENT).any(dim=1)
    assert bool((ran_out | (g["T"] <= u.T_min)).all())
    _print(f"  C2: H_eval {st['h_eval']}, H_comp {st['h_comp']}, rounds {st['rounds']}")
    # seeded 1/64 of the rays against the oracle
    rng = np.random.Generator(np.random.PCG64(5))
    idx = np.sort(rng.choice(R, R // 64, replace=False))
    r, err = _check_sample(g, o, u, W, H, idx, "C2 1/64")
    _print(f"  C2 sample: {len(idx)} rays bit-exact, max pixel error {err:.2e}, "
           f"oracle H_eval {r['stats']['h_eval']} / kernel {int((g['hit_ids'][torch.from_numpy(idx).to(DEV)] != SENT).sum())}")
    del g, first
    torch.cuda.empty_cache()


C3_CASES = [(a, k) for a in (0.01, 0.1) for k in (4, 8, 16, 32)]


@pytest.mark.parametrize("alpha,k", C3_CASES)
def test_c3_batch_workload(alpha, k):
    """64 cameras on a sphere of radius 25 looking at the origin, resolutions
    drawn from {128x128, 320x240, 640x480, 1280x720, 1920x1080} (seed 1), on
    the 1M-Gaussian scene: one batch launch captured into a hipGraph and
    replayed equals gt_render per camera (rgba, rounds and T bits), and 256
    seeded rays per camera equal the oracle."""
    _print(f"case C3 64 cameras alpha_min {alpha} k {k}")
    s, o = _c2_pair(alpha)
    s.set_tuning()
    poses, sizes = bench.cameras_c3(64)
    us = [gt.make_uniforms(gt.camera_look_at(p, [0.0, 0.0, 0.0]), w, h, hfov_deg=50.0, T_min=1e-3, alpha_min=alpha,
                           k=k) for p, (w, h) in zip(poses, sizes)]
    bufs = [dict(rgba=torch.empty((h, w, 4), dtype=torch.float32, device=DEV),
                 n_rounds=torch.empty(w * h, dtype=torch.int32, device=DEV),
                 T=torch.empty(w * h, dtype=torch.float32, device=DEV)) for w, h in sizes]
    frames = [dict(rgba=b["rgba"].data_ptr(), n_rounds=b["n_rounds"].data_ptr(), T=b["T"].data_ptr()) for b in bufs]
    batch = gt.Batch(s, us, sizes, frames)
    sp = _stream()

    def run_batch():
        for b in bufs:
            b["rgba"].fill_(-1.0)  # eager work between replays
            b["n_rounds"].fill_(-1)
        batch.render(sp, use_graph=True)
        torch.cuda.synchronize()
        return [{key: v.clone() for key, v in b.items()} for b in bufs]

    first = run_batch()   # capture + launch
    again = run_batch()   # replay
    for i, (a, b) in enumerate(zip(first, again)):
        for key in a:
            assert torch.equal(a[key].view(torch.int32), b[key].view(torch.int32)), f"camera {i}: replay {key}"
    rng = np.random.Generator(np.random.PCG64(7))
    checked = 0
    for i, (u, (w, h)) in enumerate(zip(us, sizes)):
        MR = int(first[i]["n_rounds"].max().item())
        g = _dump_render(s, u, w, h, MR)
        torch.cuda.synchronize()
        for key in ("rgba", "n_rounds", "T"):
            assert torch.equal(g[key].view(torch.int32), first[i][key].view(torch.int32)), f"camera {i}: {key}"
        idx = np.sort(rng.choice(w * h, min(256, w * h), replace=False))
        _check_sample(g, o, u, w, h, idx, f"C3 alpha {alpha} k {k} camera {i} ({w}x{h})")
        checked += len(idx)
        del g
    _print(f"  C3 alpha {alpha} k {k}: 64 cameras equal gt_render; {checked} rays bit-exact against the oracle")
    if (alpha, k) == (0.01, 16):
        tuned = gt.autotune_batch(s, batch, use_graph=True, iters=1, stream_ptr=sp)
        _print(f"  autotuner picked {tuned['best']} ({tuned['best_ms']:.1f} ms, {tuned['n_tried']} tried)")
        tuned_out = run_batch()
        for i, (a, b) in enumerate(zip(first, tuned_out)):
            for key in a:
                assert torch.equal(a[key].view(torch.int32), b[key].view(torch.int32)), f"tuned camera {i}: {key}"
        # the batch's own hit lists on its first 8 cameras, directly against the oracle
        sub = list(range(8))
        MR = max(int(first[i]["n_rounds"].max().item()) for i in sub)
        dbufs = [dict(rgba=torch.empty((sizes[i][1], sizes[i][0], 4), dtype=torch.float32, device=DEV),
                      n_rounds=torch.empty(sizes[i][0] * sizes[i][1], dtype=torch.int32, device=DEV),
                      T=torch.empty(sizes[i][0] * sizes[i][1], dtype=torch.float32, device=DEV),
                      hit_ids=torch.empty((sizes[i][0] * sizes[i][1], MR, k), dtype=torch.int32, device=DEV),
                      accept=torch.empty((sizes[i][0] * sizes[i][1], MR), dtype=torch.int32, device=DEV))
                 for i in sub]
        db = gt.Batch(s, [us[i] for i in sub], [sizes[i] for i in sub],
                      [dict(rgba=b["rgba"].data_ptr(), n_rounds=b["n_rounds"].data_ptr(), T=b["T"].data_ptr(),
                            hit_ids=b["hit_ids"].data_ptr(), accept=b["accept"].data_ptr(), max_rounds=MR)
                       for b in dbufs])
        db.render(sp, use_graph=True)
        torch.cuda.synchronize()
        for j, i in enumerate(sub):
            w, h = sizes[i]
            idx = np.sort(rng.choice(w * h, min(256, w * h), replace=False))
            _check_sample(dbufs[j], o, us[i], w, h, idx, f"C3 batch dump camera {i}")
        db.close()
        s.set_tuning()
    batch.close()
    del bufs, first, again
    torch.cuda.empty_cache()


C5_PICKS = {"mean": (slice(0, 3), 96), "f_dc": (slice(3, 6), 64), "f_rest": (slice(6, 51), 96),
            "logit": (slice(51, 52), 64), "log_scale": (slice(52, 55), 96), "quat": (slice(55, 59), 96)}


def test_c5_gradients_fd_and_modes():
    """3,000,000 Gaussians (256-cluster mixture in a 40^3 box, log-scales
    N(-4.5, 0.7)), 16 cameras at 1280x720 on a sphere of radius 25 (seed 1),
    alpha_min 0.01, k 16, T_min 0.001; loss = sum over views of
    G_v . render_v with G_v = 2 (render_v - target_v), the target rendered
    from means + N(0, 0.01) (seed 3).  Checkpoint-mode gradients (accumulated
    across the views with FP32 atomics) equal recompute-mode gradients within
    FP32-atomic tolerance, and match the FP64 reference on 512 seeded
    parameters (seed 2): autograd of the FP64 re-render of the frozen hit
    lists over the rays that composite each picked Gaussian, within
    1e-3 relative.  Central differences confirm the reference: with step
    1e-4 alone their O(h^2) error reaches 2e-3 relative for C5's smallest
    Gaussians (scales near 0.004, so h is 2.5 % of sigma), so the check uses
    steps 1e-4 and 5e-5 with Richardson extrapolation (4 f(h/2) - f(h)) / 3."""
    _print("case C5 3M 16 x 1280x720 k16: checkpoint vs recompute, FD on 512 parameters")
    for key in [k for k in _CACHE if k != "raw"]:
        _CACHE.pop(key)[0].close()  # the 1M scenes are no longer needed
    raw = gt.generate_raw(seed=0, config="c5")
    n = raw.shape[0]
    s = gt.Scene(raw, alpha_min=0.01)
    W, H, V, cap = 1280, 720, 16, 256
    R = W * H
    rngc = np.random.Generator(np.random.PCG64(1))
    dirs = rngc.normal(size=(V, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    us = [gt.make_uniforms(gt.camera_look_at(list(25.0 * d), [0.0, 0.0, 0.0]), W, H, hfov_deg=50.0, T_min=1e-3,
                           alpha_min=0.01, k=16) for d in dirs]
    sp = _stream()
    rng = np.random.Generator(np.random.PCG64(3))
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    targets = torch.empty((V, H, W, 4), dtype=torch.float32, device=DEV)
    ts = gt.Scene(raw_t, alpha_min=0.01)
    for v, u in enumerate(us):
        ts.render_device(u, W, H, targets[v].data_ptr(), sp)
    torch.cuda.synchronize()
    ts.close()
    del raw_t
    _print("  C5: scene and 16 targets ready")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=DEV)
    cids = torch.empty((R, cap), dtype=torch.int32, device=DEV)
    ncomp = torch.empty(R, dtype=torch.int32, device=DEV)
    ovf = torch.zeros(1, dtype=torch.int32, device=DEV)
    G = torch.empty((V, H, W, 4), dtype=torch.float32, device=DEV)
    # the picks: seeded Gaussians among those view 0 composites
    s.render_device(us[0], W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(), n_comp_ptr=ncomp.data_ptr(),
                    comp_cap=cap)
    valid = torch.arange(cap, device=DEV)[None, :] < ncomp[:, None].clamp(max=cap)
    seen = torch.unique(cids[valid]).cpu().numpy()
    pick_rng = np.random.default_rng(2)
    picks = []
    for name, (sl, cnt) in C5_PICKS.items():
        picks += [(int(pick_rng.choice(seen)), int(pick_rng.integers(sl.start, sl.stop))) for _ in range(cnt)]
    picked = torch.from_numpy(np.unique([g for g, _ in picks]).astype(np.int32)).to(DEV)
    # checkpoint mode, gathering the rows of rays that composite a picked Gaussian
    grads_ck = torch.zeros((n, gt.GT_RAW_FLOATS), dtype=torch.float32, device=DEV)
    rows = []
    most = 0
    for v, u in enumerate(us):
        s.render_device(u, W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(), n_comp_ptr=ncomp.data_ptr(),
                        comp_cap=cap)
        torch.sub(rgba, targets[v], out=G[v])
        G[v].mul_(2.0)
        gt.backward(s, u, W, H, G[v].data_ptr(), grads_ck.data_ptr(), cids.data_ptr(), ncomp.data_ptr(), cap,
                    ovf.data_ptr(), 0, sp)
        most = max(most, int(ncomp.max().item()))
        valid = torch.arange(cap, device=DEV)[None, :] < ncomp[:, None].clamp(max=cap)
        has = (torch.isin(cids, picked) & valid).any(dim=1)
        r_idx = torch.nonzero(has).flatten()
        rows.append(dict(pix=r_idx.cpu().numpy(), cids=cids[r_idx].cpu().numpy().view(np.uint32),
                         ncomp=ncomp[r_idx].cpu().numpy(), G=G[v].reshape(R, 4)[r_idx, :3].cpu().numpy()))
    torch.cuda.synchronize()
    assert int(ovf.item()) == 0 and most <= cap, f"checkpoint overflow (most {most})"
    # recompute mode from the same dL/dRGBA
    grads_rc = torch.zeros_like(grads_ck)
    for v, u in enumerate(us):
        gt.backward(s, u, W, H, G[v].data_ptr(), grads_rc.data_ptr(), None, None, 0, ovf.data_ptr(), 0, sp)
    torch.cuda.synchronize()
    assert int(ovf.item()) == 0
    g_ck = grads_ck.cpu().numpy().astype(np.float64)
    g_rc = grads_rc.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g_rc - g_ck) / np.linalg.norm(g_ck)
    _print(f"  C5 modes: relative L2 {rel:.2e}, max composited per ray {most}")
    assert rel < 1e-5
    np.testing.assert_allclose(g_rc, g_ck, rtol=1e-4, atol=1e-6 * np.abs(g_ck).max())
    s.close()
    del grads_ck, grads_rc, G, targets, cids
    torch.cuda.empty_cache()
    # FD of the FP64 loss sum_v G_v . render64_v over the rays compositing each picked Gaussian
    raw64 = torch.from_numpy(raw.astype(np.float64))
    del raw
    per_view = []
    for v, u in enumerate(us):
        rw = rows[v]
        rr, gg, pp = fr.flatten_checkpoint(rw["cids"], rw["ncomp"])
        order = np.argsort(gg, kind="stable")
        p = orc.params_from_uniforms(u, W, H)
        d = np.stack([orc.ray_dir(p, int(x) % W, int(x) // W) for x in rw["pix"]]).astype(np.float64) \
            if len(rw["pix"]) else np.zeros((0, 3))
        per_view.append(dict(gs=gg[order], rs=rr[order], dirs=d,
                             o=torch.tensor([u.to_world[12], u.to_world[13], u.to_world[14]], dtype=torch.float64)))
    # per pick: the FP64 loss over only the Gaussians those rays composite
    # (a small leaf tensor, so autograd and the differences stay cheap)
    h = 1e-4
    ag = np.zeros(len(picks))
    fd = np.zeros(len(picks))
    fd_plain = np.zeros(len(picks))
    for i, (gid, col) in enumerate(picks):
        terms, used = [], []
        for v, pv in enumerate(per_view):
            lo, hi = np.searchsorted(pv["gs"], gid), np.searchsorted(pv["gs"], gid, side="right")
            if lo == hi:
                continue
            sub = np.unique(pv["rs"][lo:hi])
            srays, sgids, spos = fr.flatten_checkpoint(rows[v]["cids"][sub], rows[v]["ncomp"][sub])
            terms.append([pv["o"], torch.from_numpy(pv["dirs"][sub]), srays, sgids, spos, len(sub),
                          torch.from_numpy(rows[v]["G"][sub].astype(np.float64))])
            used.append(sgids)
        uniq = np.unique(np.concatenate(used))
        for t in terms:
            t[3] = np.searchsorted(uniq, t[3])
        row = int(np.searchsorted(uniq, gid))
        sub_raw = raw64[torch.from_numpy(uniq)].clone()

        def loss(x):
            tot = 0.0
            for o_, d_, sr, sg, spo, m, Gs in terms:
                tot = tot + (fr.render(x, o_, d_, sr, sg, spo, m) * Gs).sum()
            return tot

        x = sub_raw.clone().requires_grad_(True)
        loss(x).backward()
        ag[i] = float(x.grad[row, col])
        with torch.no_grad():
            def central(step):
                x2 = sub_raw.clone()
                x2[row, col] += step
                lp = float(loss(x2))
                x2[row, col] -= 2 * step
                lm = float(loss(x2))
                return (lp - lm) / (2 * step)
            f1, f2 = central(h), central(h / 2)
            fd_plain[i] = f1
            fd[i] = (4.0 * f2 - f1) / 3.0  # Richardson: cancels the O(h^2) truncation term
    gv = np.array([g_ck[a, b] for a, b in picks])
    scale = np.abs(ag).max()
    tol = 1e-3 * (np.abs(ag) + 1e-2 * scale)
    err = np.abs(gv - ag)
    err_fd = np.abs(fd - ag)
    _print(f"  C5 vs FP64 autograd: max |kernel - ref| {err.max():.3e}, max |ref| {scale:.3e}, "
           f"worst ratio {(err / tol).max():.3f}; Richardson FD vs autograd worst ratio {(err_fd / tol).max():.3f}; "
           f"plain step-1e-4 FD worst ratio {(np.abs(fd_plain - ag) / tol).max():.3f}")
    bad = np.nonzero(err_fd > tol)[0]
    assert bad.size == 0, f"FD does not confirm the FP64 reference on {bad.size} picks: {[picks[i] for i in bad[:4]]}"
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, f"{bad.size} of 512 differ, first {[picks[i] for i in bad[:4]]} {gv[bad[:4]]} {ag[bad[:4]]}"
