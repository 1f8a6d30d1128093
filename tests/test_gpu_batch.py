"""GPU: batched multi-camera rendering and launch tuning never change results.

Bar: each camera of a batch yields the same hit lists, accept masks, round
counts and pixels (bit for bit) as gt_render on that camera alone; every
gt_tuning setting (tile kernel or persistent kernel, ray-tile shape, LDS
stack depth, scheduling thresholds) gives bit-identical outputs; hipGraph
replays equal eager launches.
"""
import os

import numpy as np
import pytest
import torch

import gtrace as gt
from helpers import uniforms

pytestmark = pytest.mark.gpu

SIZES = [(37, 23), (64, 48), (128, 96), (33, 65), (96, 64), (8, 8)]


def _scene():
    raw = gt.generate_raw(n=60000, seed=21, config="c2", n_clusters=6)
    s = gt.Scene(raw, alpha_min=0.01)
    c = raw[0::6, :3].mean(axis=0)
    return s, c


def _cams(c, k=8):
    rng = np.random.default_rng(5)
    out = []
    for i, (w, h) in enumerate(SIZES):
        off = rng.normal(size=3)
        off = off / np.linalg.norm(off) * 2.5
        out.append(uniforms(list(c + off), list(c), w, h, hfov_deg=55, alpha_min=0.01, k=k, T_min=1e-3))
    return out


def _render_batch(s, us, max_rounds, use_graph=False, replays=1):
    dev = torch.device("cuda:0")
    bufs = []
    frames = []
    for u, (w, h) in zip(us, SIZES):
        b = dict(rgba=torch.zeros((h, w, 4), device=dev),
                 hits=torch.zeros((w * h, max_rounds, u.k), dtype=torch.int32, device=dev),
                 acc=torch.zeros((w * h, max_rounds), dtype=torch.int32, device=dev),
                 rounds=torch.zeros(w * h, dtype=torch.int32, device=dev))
        bufs.append(b)
        frames.append(dict(rgba=b["rgba"].data_ptr(), hit_ids=b["hits"].data_ptr(), accept=b["acc"].data_ptr(),
                           n_rounds=b["rounds"].data_ptr(), max_rounds=max_rounds))
    batch = gt.Batch(s, us, SIZES, frames)
    st = torch.cuda.current_stream().cuda_stream
    outs = []
    for r in range(replays):
        for b in bufs:
            b["rgba"].fill_(-1.0)  # eager work between replays
        batch.render(st, use_graph=use_graph)
        torch.cuda.synchronize()
        outs.append([{k: v.cpu().numpy().copy() for k, v in b.items()} for b in bufs])
    batch.close()
    return outs


def _single(s, us, max_rounds):
    res = []
    for u, (w, h) in zip(us, SIZES):
        res.append(s.render(u, w, h, max_rounds=max_rounds))
    return res


def _same(batch_out, single):
    for i, (b, r) in enumerate(zip(batch_out, single)):
        assert np.array_equal(b["hits"].view(np.uint32), r["hit_ids"]), f"camera {i}: hit lists"
        assert np.array_equal(b["acc"].view(np.uint32), r["accept"]), f"camera {i}: accept"
        assert np.array_equal(b["rounds"].view(np.uint32), r["n_rounds"]), f"camera {i}: rounds"
        assert b["rgba"].tobytes() == r["rgba"].tobytes(), f"camera {i}: pixels"


@pytest.mark.parametrize("path", [1, 2, 3, 4])
def test_batch_equals_single_camera(path):
    print(f"case batch path {path}", flush=True)
    s, c = _scene()
    us = _cams(c)
    s.set_tuning(path=path)
    single = _single(s, us, 32)
    out = _render_batch(s, us, 32)[0]
    _same(out, single)


@pytest.mark.parametrize("path", [1, 2, 3, 4])
def test_batch_graph_replays(path):
    print(f"case batch graph path {path}", flush=True)
    s, c = _scene()
    us = _cams(c)
    s.set_tuning(path=path)
    single = _single(s, us, 32)
    outs = _render_batch(s, us, 32, use_graph=True, replays=3)
    for o in outs:
        _same(o, single)


TUNINGS = [dict(path=1, tile_w=8), dict(path=1, tile_w=16), dict(path=1, tile_w=32, stack_lds=8),
           dict(path=1, tile_w=4, stack_lds=64), dict(path=2, refill_min=1, comp_min=1),
           dict(path=2, refill_min=64, comp_min=64, tile_w=16, stack_lds=16), dict(path=2, blocks_per_cu=1),
           dict(path=3, tile_w=8), dict(path=3, tile_w=4), dict(path=3, tile_w=32),
           dict(path=4, tile_w=8), dict(path=4, tile_w=4, blocks_per_cu=1), dict(path=4, tile_w=16, blocks_per_cu=8)]


def test_tuning_never_changes_results():
    print("case tuning invariance", flush=True)
    s, c = _scene()
    u = _cams(c, k=16)[2]
    w, h = SIZES[2]
    ref = None
    for t in TUNINGS:
        s.set_tuning(**t)
        r = s.render(u, w, h, max_rounds=32)
        if ref is None:
            ref = r
            continue
        for key in ("hit_ids", "accept", "n_rounds"):
            assert np.array_equal(r[key], ref[key]), (t, key)
        assert r["rgba"].tobytes() == ref["rgba"].tobytes(), t


@pytest.mark.parametrize("k,alpha", [(4, 0.1), (32, 0.1), (16, 0.01)])
def test_batch_k_and_alpha_of_the_batched_workload(k, alpha):
    """The batched workload's settings: alpha_min in {0.01, 0.1} (scene filter
    and uniform) and k from {4, 8, 16, 32}; a batch shares one k."""
    print(f"case batch k={k} alpha={alpha}", flush=True)
    raw = gt.generate_raw(n=60000, seed=21, config="c2", n_clusters=6)
    s = gt.Scene(raw, alpha_min=alpha)
    c = raw[0::6, :3].mean(axis=0)
    rng = np.random.default_rng(7)
    us = []
    for w, h in SIZES:
        off = rng.normal(size=3)
        off = off / np.linalg.norm(off) * 2.5
        us.append(uniforms(list(c + off), list(c), w, h, hfov_deg=55, alpha_min=alpha, k=k, T_min=1e-3))
    single = _single(s, us, 48)
    _same(_render_batch(s, us, 48)[0], single)


def test_batch_rejects_mixed_k():
    s, c = _scene()
    us = _cams(c, k=8)
    us[1] = uniforms(list(c + 2.5), list(c), SIZES[1][0], SIZES[1][1], hfov_deg=55, alpha_min=0.01, k=16,
                     T_min=1e-3)
    dev = torch.device("cuda:0")
    imgs = [torch.zeros((h, w, 4), device=dev) for (w, h) in SIZES]
    frames = [dict(rgba=t.data_ptr()) for t in imgs]
    with pytest.raises(gt.GTError, match="shares k"):
        gt.Batch(s, us, SIZES, frames)


def test_graph_replay_after_scene_rebuild():
    """A captured batch graph holds the scene's device arrays by value.  After
    a rebuild moves them (gt_scene_set_alpha_min; a refit whose proxy set
    changed, which rebuilds on the GPU) the next replay must render the new
    scene: equal, camera by camera, to gt_render."""
    print("case batch graph after rebuilds", flush=True)
    s, c = _scene()
    us = _cams(c)
    MR = 64
    _same(_render_batch(s, us, MR, use_graph=True, replays=1)[0], _single(s, us, MR))
    dev = torch.device("cuda:0")
    frames, bufs = [], []
    for u, (w, h) in zip(us, SIZES):
        b = dict(rgba=torch.zeros((h, w, 4), device=dev),
                 hits=torch.zeros((w * h, MR, u.k), dtype=torch.int32, device=dev),
                 acc=torch.zeros((w * h, MR), dtype=torch.int32, device=dev),
                 rounds=torch.zeros(w * h, dtype=torch.int32, device=dev))
        bufs.append(b)
        frames.append(dict(rgba=b["rgba"].data_ptr(), hit_ids=b["hits"].data_ptr(), accept=b["acc"].data_ptr(),
                           n_rounds=b["rounds"].data_ptr(), max_rounds=MR))
    batch = gt.Batch(s, us, SIZES, frames)
    st = torch.cuda.current_stream().cuda_stream

    def replay():
        for b in bufs:
            b["rgba"].fill_(-1.0)
        batch.render(st, use_graph=True)
        torch.cuda.synchronize()
        return [{k: v.cpu().numpy().copy() for k, v in b.items()} for b in bufs]

    _same(replay(), _single(s, us, MR))
    s.set_alpha_min(0.05)  # re-filter: every device array is rebuilt
    _same(replay(), _single(s, us, MR))
    raw = gt.generate_raw(n=60000, seed=21, config="c2", n_clusters=6)
    raw2 = raw.copy()
    raw2[::7, 51] = -9.0  # opacity far below alpha_min: the proxy set changes
    s2 = gt.Scene(raw, alpha_min=0.05)
    b2 = gt.Batch(s2, us, SIZES, frames)
    b2.render(st, use_graph=True)
    torch.cuda.synchronize()
    x = torch.from_numpy(raw2).to(dev)
    assert s2.refit_device(x.data_ptr(), st) is True  # fell back to a GPU rebuild
    for b in bufs:
        b["rgba"].fill_(-1.0)
    b2.render(st, use_graph=True)
    torch.cuda.synchronize()
    _same([{k: v.cpu().numpy().copy() for k, v in b.items()} for b in bufs], _single(s2, us, MR))
    s3 = gt.Scene(raw2, alpha_min=0.05)
    _same([{k: v.cpu().numpy().copy() for k, v in b.items()} for b in bufs], _single(s3, us, MR))
    b2.close()
    batch.close()
    for sc in (s, s2, s3):
        sc.close()


def test_first_render_of_a_process_is_a_graph_batch():
    """The face table is uploaded at scene creation, outside any capture: a
    process whose first render is a captured batch renders correctly."""
    import subprocess
    import sys
    print("case graph batch as the first render of a process", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    code = (
        "import sys; sys.path[:0] = %r\n"
        "import numpy as np, torch, gtrace as gt\n"
        "from test_gpu_batch import _scene, _cams, _render_batch, _single, _same\n"
        "s, c = _scene(); us = _cams(c)\n"
        "out = _render_batch(s, us, 64, use_graph=True, replays=2)\n"
        "for o in out: _same(o, _single(s, us, 64))\n"
        "print('first-render graph batch ok', flush=True)\n"
    ) % ([here, os.path.join(os.path.dirname(here), "webgpu-gaussiantracer_amd"),
          os.path.join(os.path.dirname(here), "oracle")],)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "first-render graph batch ok" in r.stdout


def test_one_scene_on_two_streams():
    """Renders of one scene issued on two streams overlap on the GPU.  The
    persistent kernel's queue counters are per stream, so every image equals
    its render on the default stream alone, bit for bit."""
    print("case one scene, two streams", flush=True)
    s, c = _scene()
    us = _cams(c, k=16)[:2]
    sizes = SIZES[:2]
    dev = torch.device("cuda:0")
    ref = [s.render(u, w, h)["rgba"] for u, (w, h) in zip(us, sizes)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    imgs = [[torch.empty((h, w, 4), device=dev) for _ in range(6)] for (w, h) in sizes]
    torch.cuda.synchronize()
    for rep in range(6):
        for i, (u, (w, h)) in enumerate(zip(us, sizes)):
            with torch.cuda.stream(streams[i]):
                s.render_device(u, w, h, imgs[i][rep].data_ptr(), streams[i].cuda_stream)
    torch.cuda.synchronize()
    for i in range(2):
        for rep in range(6):
            assert imgs[i][rep].cpu().numpy().tobytes() == ref[i].tobytes(), f"stream {i} render {rep}"
    s.close()


def test_autotune_frame_keeps_results_and_stores_its_choice():
    """gt_autotune_frame times the launch candidates on the caller's own
    frame and stores the fastest in the scene: the frame rendered after it
    (with the stored tuning) equals the frame rendered before it bit for bit,
    the scene reports the chosen tuning, and the times it returns are
    those of real launches."""
    print("case autotune_frame", flush=True)
    s, c = _scene()
    u = _cams(c, k=16)[2]
    w, h = SIZES[2]
    before = s.render(u, w, h, max_rounds=32)
    dev = torch.device("cuda:0")
    img = torch.zeros((h, w, 4), device=dev)
    res = gt.autotune_frame(s, u, w, h, img.data_ptr(), iters=2,
                            stream_ptr=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert res["n_tried"] >= 5 and 0.0 < res["best_ms"] <= res["first_ms"]
    assert s.get_tuning() == res["best"]
    assert img.cpu().numpy().tobytes() == before["rgba"].tobytes()  # its frames are the caller's frame
    after = s.render(u, w, h, max_rounds=32)
    for key in ("hit_ids", "accept", "n_rounds"):
        assert np.array_equal(after[key], before[key]), key
    assert after["rgba"].tobytes() == before["rgba"].tobytes()
    s.close()
