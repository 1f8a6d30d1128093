"""Scene refit (gt_scene_refit_device): new parameters, same tree.

Bar: after a refit, renders equal those of a fresh build of the same
parameters -- hit lists, accept masks, round counts, T and RGBA bit for bit --
for host-built (binned SAH, depth-first node order) and GPU-built (Morton
radix tree, level order) trees.  A refit with unchanged parameters leaves the
tree exactly as built: the traversal work counters (nodes, proxies, hits),
which depend on every node box, are unchanged too.  A change of the proxy set
falls back to a GPU rebuild.  Training with refits equals training with
rebuilds bit for bit.
"""
import numpy as np
import pytest
import torch

from helpers import uniforms
import gtrace as gt
from gtrace.train import Trainer

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _render(s, u, W, H):
    return s.render(u, W, H, max_rounds=64, want_stats=True)


def _assert_same(ga, gb, label, work=False):
    for key in ("hit_ids", "accept", "n_rounds"):
        np.testing.assert_array_equal(ga[key], gb[key], err_msg=f"{label}: {key}")
    np.testing.assert_array_equal(ga["T"].view(np.uint32), gb["T"].view(np.uint32), err_msg=f"{label}: T")
    np.testing.assert_array_equal(ga["rgba"].view(np.uint32), gb["rgba"].view(np.uint32), err_msg=f"{label}: rgba")
    keys = ("rays", "rounds", "h_eval", "h_comp") + (("nodes", "proxies", "hits") if work else ())
    for key in keys:
        assert ga["stats"][key] == gb["stats"][key], (label, key, ga["stats"][key], gb["stats"][key])


def _scene(raw, alpha, kind):
    if kind == "host":
        return gt.Scene(raw, alpha_min=alpha)
    return gt.Scene(raw, alpha_min=alpha, gpu_build=True)


CASES = [("c1", 3000, 0.1, [0, 0, 8]), ("c4", 20000, 0.2, [0, 0, 6]), ("c2", 60000, 0.01, [0.3, 0.2, 0.1])]


@pytest.mark.parametrize("kind", ["host", "gpu"])
@pytest.mark.parametrize("cfg,n,alpha,eye", CASES)
def test_refit_unchanged_keeps_tree(kind, cfg, n, alpha, eye):
    _print(f"case refit identity {kind} {cfg} n={n}")
    raw = gt.generate_raw(n=n, seed=3, config=cfg)
    s = _scene(raw, alpha, kind)
    u = uniforms(eye, [0, 0, 0], 160, 120, hfov_deg=60, alpha_min=alpha, k=8, T_min=0.01)
    before = _render(s, u, 160, 120)
    info0 = s.info()
    t = torch.from_numpy(raw).to("cuda:0")
    assert s.refit_device(t.data_ptr()) is False
    torch.cuda.synchronize()
    after = _render(s, u, 160, 120)
    info1 = s.info()
    for key in ("n_proxies", "n_nodes", "max_stack", "bvh_depth"):
        assert info0[key] == info1[key], key
    _assert_same(before, after, f"identity {kind} {cfg}", work=True)
    # the downloaded records equal a fresh host build's
    h = gt.Scene(raw, alpha_min=alpha)
    hd, sd = h.host_data(), s.host_data()
    for key in ("packed", "o2w", "w2o"):
        np.testing.assert_array_equal(hd[key].view(np.uint32), sd[key].view(np.uint32), err_msg=key)
    np.testing.assert_array_equal(hd["proxy_ids"], sd["proxy_ids"])


@pytest.mark.parametrize("kind", ["host", "gpu"])
@pytest.mark.parametrize("cfg,n,alpha,eye", CASES)
def test_refit_moved_equals_fresh_build(kind, cfg, n, alpha, eye):
    """Means, scales, rotations and colours move (opacities kept): three refits in a row."""
    _print(f"case refit moved {kind} {cfg} n={n}")
    raw = gt.generate_raw(n=n, seed=4, config=cfg)
    s = _scene(raw, alpha, kind)
    rng = np.random.default_rng(11)
    t = torch.from_numpy(raw).to("cuda:0")
    u = uniforms(eye, [0, 0, 0], 160, 120, hfov_deg=60, alpha_min=alpha, k=8, T_min=0.01)
    cur = raw.copy()
    for step in range(3):
        cur = cur.copy()
        cur[:, 0:3] += rng.normal(0, 0.02, (n, 3)).astype(np.float32)
        cur[:, 3:51] += rng.normal(0, 0.05, (n, 48)).astype(np.float32)
        cur[:, 52:55] += rng.normal(0, 0.05, (n, 3)).astype(np.float32)
        cur[:, 55:59] += rng.normal(0, 0.05, (n, 4)).astype(np.float32)
        t.copy_(torch.from_numpy(cur))
        assert s.refit_device(t.data_ptr()) is False, step
        torch.cuda.synchronize()
        fresh = gt.Scene(cur, alpha_min=alpha)
        _assert_same(_render(fresh, u, 160, 120), _render(s, u, 160, 120), f"moved {kind} {cfg} step {step}")


@pytest.mark.parametrize("kind", ["host", "gpu"])
def test_refit_proxy_set_change_rebuilds(kind):
    _print(f"case refit proxy change {kind}")
    n, alpha = 8000, 0.2
    raw = gt.generate_raw(n=n, seed=6, config="c4")
    s = _scene(raw, alpha, kind)
    t = torch.from_numpy(raw).to("cuda:0")
    u = uniforms([0, 0, 6], [0, 0, 0], 128, 96, hfov_deg=60, alpha_min=alpha, k=8, T_min=0.01)
    op = 1.0 / (1.0 + np.exp(-raw[:, 51].astype(np.float64)))
    lose = int(np.flatnonzero(op > alpha)[0])
    gain = int(np.flatnonzero(op <= alpha)[0])
    for label, idx, logit in (("lose", lose, -8.0), ("gain", gain, 4.0)):
        cur = raw.copy()
        cur[idx, 51] = logit
        t.copy_(torch.from_numpy(cur))
        assert s.refit_device(t.data_ptr()) is True, label
        torch.cuda.synchronize()
        fresh = gt.Scene(cur, alpha_min=alpha)
        assert s.info()["n_proxies"] == fresh.info()["n_proxies"]
        _assert_same(_render(fresh, u, 128, 96), _render(s, u, 128, 96), f"{label} {kind}")
        # and the rebuilt scene refits again
        assert s.refit_device(t.data_ptr()) is False
        torch.cuda.synchronize()
        _assert_same(_render(fresh, u, 128, 96), _render(s, u, 128, 96), f"{label} {kind} refit after rebuild")


def test_refit_then_alpha_change():
    """A refit scene holds device data only; set_alpha_min rebuilds it on the GPU."""
    _print("case refit then alpha")
    raw = gt.generate_raw(n=6000, seed=8, config="c4")
    s = gt.Scene(raw, alpha_min=0.2)
    t = torch.from_numpy(raw).to("cuda:0")
    s.refit_device(t.data_ptr())
    s.set_alpha_min(0.3)
    u = uniforms([0, 0, 6], [0, 0, 0], 128, 96, hfov_deg=60, alpha_min=0.3, k=8, T_min=0.01)
    _assert_same(_render(gt.Scene(raw, alpha_min=0.3), u, 128, 96), _render(s, u, 128, 96), "alpha after refit")


def test_trainer_refit_scene_tracks_parameters():
    """Refit-based training: after every step the scene renders exactly as a
    fresh build of the current parameters.  (Gradients are accumulated with
    float atomics, so two trainers agree to rounding, not bit for bit.)"""
    _print("case trainer refit")
    n = 4000
    raw0 = gt.generate_raw(n=n, seed=9, config="c4")
    dev = torch.device("cuda:0")
    W, H = 96, 80
    views = []
    for eye in ([0, 0, 6], [5, 1, 2]):
        u = uniforms(eye, [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.05, k=8, T_min=0.01)
        tgt = gt.Scene(raw0, alpha_min=0.05).render(u, W, H)["rgba"]
        views.append((u, W, H, torch.from_numpy(tgt).to(dev)))
    start = raw0.copy()
    start[:, 0:3] += np.random.default_rng(2).normal(0, 0.05, (n, 3)).astype(np.float32)
    runs = {}
    for build, refit in (("host", True), ("gpu", True), ("gpu", False)):
        raw = torch.from_numpy(start).to(dev)
        tr = Trainer(raw, views, alpha_min=0.05, build=build, refit=refit)
        losses = []
        for step in range(4):
            losses.append(tr.step())
            torch.cuda.synchronize()
            cur = raw.cpu().numpy()
            fresh = gt.Scene(cur, alpha_min=0.05)
            u0 = views[0][0]
            _assert_same(_render(fresh, u0, W, H), _render(tr.scene, u0, W, H), f"trainer {build} {refit} step {step}")
        runs[(build, refit)] = losses
        assert losses[-1] < losses[0], (build, refit, losses)
        tr.close()
    ref = runs[("gpu", False)]
    for key, losses in runs.items():
        np.testing.assert_allclose(losses, ref, rtol=1e-3, err_msg=str(key))
