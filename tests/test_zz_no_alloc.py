"""GPU: the allocation-free batched render (gt_batch_render_noalloc) against
the allocating one (gt_batch_create + gt_batch_render), through the C-ABI.

Bar: bit-identical outputs per camera (pixels, hit lists, accept masks,
round counts, T), on two calls in a row where the second renders fewer and
smaller cameras (the shared device array keeps the first call's table), for
the packet tile kernel and the persistent kernel (whose work counter lives
in the same array); more cameras than the array holds is GT_E_LIMIT; and
a call after the first on a device leaves its free memory unchanged.
"""
import numpy as np
import pytest
import torch

import gtrace as gt
from helpers import uniforms

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _scene():
    raw = gt.generate_raw(n=60000, seed=21, config="c2", n_clusters=6)
    s = gt.Scene(raw, alpha_min=0.01)
    return s, raw[0::6, :3].mean(axis=0)


def _cams(c, sizes, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for w, h in sizes:
        off = rng.normal(size=3)
        off = off / np.linalg.norm(off) * 2.5
        out.append(uniforms(list(c + off), list(c), w, h, hfov_deg=55, alpha_min=0.01, k=k, T_min=1e-3))
    return out


def _buffers(us, sizes, MR):
    bufs, frames = [], []
    for u, (w, h) in zip(us, sizes):
        b = dict(rgba=torch.full((h, w, 4), -1.0, device=DEV),
                 hits=torch.zeros((w * h, MR, u.k), dtype=torch.int32, device=DEV),
                 acc=torch.zeros((w * h, MR), dtype=torch.int32, device=DEV),
                 rounds=torch.zeros(w * h, dtype=torch.int32, device=DEV),
                 T=torch.zeros(w * h, dtype=torch.float32, device=DEV))
        bufs.append(b)
        frames.append(dict(rgba=b["rgba"].data_ptr(), hit_ids=b["hits"].data_ptr(), accept=b["acc"].data_ptr(),
                           n_rounds=b["rounds"].data_ptr(), T=b["T"].data_ptr(), max_rounds=MR))
    return bufs, frames


def _host(bufs):
    return [{k: v.cpu().numpy().copy() for k, v in b.items()} for b in bufs]


def _equal(a, b, label):
    for i, (x, y) in enumerate(zip(a, b)):
        for key in x:
            assert x[key].tobytes() == y[key].tobytes(), f"{label}: camera {i} {key}"


@pytest.mark.parametrize("path", [3, 2])
def test_noalloc_batch_equals_allocating_batch(path):
    print(f"case no-alloc batch, tuning path {path}", flush=True)
    s, c = _scene()
    s.set_tuning(path=path)
    st = torch.cuda.current_stream().cuda_stream
    MR = 64
    calls = [([(37, 23), (64, 48), (128, 96), (33, 65), (96, 64), (8, 8)], 8, 5),
             ([(20, 11), (40, 30), (9, 17)], 8, 6)]
    free_after = []
    for j, (sizes, k, seed) in enumerate(calls):
        us = _cams(c, sizes, k, seed)
        ref_bufs, ref_frames = _buffers(us, sizes, MR)
        b = gt.Batch(s, us, sizes, ref_frames)
        b.render(st)
        torch.cuda.synchronize()
        b.close()
        ref = _host(ref_bufs)
        bufs, frames = _buffers(us, sizes, MR)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        gt.batch_render_noalloc(s, us, sizes, frames, stream_ptr=st)
        torch.cuda.synchronize()
        free_after.append((free0, torch.cuda.mem_get_info()[0]))
        _equal(_host(bufs), ref, f"call {j + 1}")
        # each camera also equals gt_render on that camera alone
        for i, (u, (w, h)) in enumerate(zip(us, sizes)):
            one = s.render(u, w, h, max_rounds=MR)
            assert one["rgba"].tobytes() == ref[i]["rgba"].tobytes(), f"call {j + 1} camera {i}: gt_render"
    # the first call on a device makes the runtime load the library's device
    # code holding the array (code objects load on first use); from then on
    # a call leaves the device's free memory as it was
    a, b_ = free_after[1]
    assert a == b_, f"call 2: device free memory changed from {a} to {b_}"
    s.close()


def test_noalloc_limits():
    print("case no-alloc limits", flush=True)
    s, c = _scene()
    n = gt.NOALLOC_MAX_CAMERAS + 1
    us = _cams(c, [(8, 8)] * n, 8, 7)
    img = torch.zeros((8, 8, 4), device=DEV)
    frames = [dict(rgba=img.data_ptr())] * n
    with pytest.raises(gt.GTError) as e:
        gt.batch_render_noalloc(s, us, [(8, 8)] * n, frames)
    assert e.value.code == -5 and "256" in str(e.value)
    us2 = _cams(c, [(8, 8), (8, 8)], 8, 8)
    us2[1].k = 4
    with pytest.raises(gt.GTError) as e:
        gt.batch_render_noalloc(s, us2, [(8, 8)] * 2, frames[:2])
    assert e.value.code == -1
    s.close()
