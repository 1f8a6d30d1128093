"""The reference's antialiasing path (GS_shader.ts:6-31, 281-307), compiled
out there (antialiasing = false, GS_shader.ts:54) and opt-in here.

Bar: with jittered rays, hit lists, accept masks, round counts and T are
bit-exact against the oracle (which restates getNewSeed/rng and the running
average); each frame's pixels are within 1e-4 of the FP64 re-evaluation of
its hit lists, averaged as the shader averages; the device and host paths
agree bit for bit.
"""
import numpy as np
import pytest
import torch

from helpers import PIXEL_TOL, uniforms
import gtrace as gt
import oracle as orc

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _scene():
    """The scene renders with screen-rectangle culling and precomputed
    colours forced on (tuning rects=1, colors=1): the rectangles must hold
    every jittered ray, and the colours of the camera origin serve every
    jittered ray (the origin does not move)."""
    raw = gt.generate_raw(n=20000, seed=0, config="c1")
    s = gt.Scene(raw, alpha_min=0.1)
    s.set_tuning(rects=1, colors=1)
    return raw, s, orc.OracleScene(raw, 0.1)


def test_antialias_frames_parity():
    _print("case antialias 4 frames 96x72 k6")
    raw, s, o = _scene()
    W, H = 96, 72
    prev = None
    imgs = []
    for frames in range(4):
        u = uniforms([0, 0, 8], [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03,
                     accumulated_frames=frames)
        g = s.render(u, W, H, max_rounds=32, want_stats=True, antialiasing=True, rgba_prev=prev)
        p = orc.params_from_uniforms(u, W, H, antialias=True)
        r = orc.render(o, p, max_rounds=32, rgba_prev=None if prev is None else prev.reshape(-1, 4))
        np.testing.assert_array_equal(g["n_rounds"], r["n_rounds"])
        np.testing.assert_array_equal(g["hit_ids"], r["hit_ids"])
        np.testing.assert_array_equal(g["accept"], r["accept"])
        np.testing.assert_array_equal(g["T"].view(np.uint32), r["T"].view(np.uint32))
        rgb64, _ = orc.reeval_f64(o, p, g["n_rounds"], g["hit_ids"], g["accept"])
        if frames == 0:
            want = rgb64
        else:
            a = 1.0 / (frames + 1.0)
            want = prev.reshape(-1, 4)[:, :3].astype(np.float64) * (1.0 - a) + rgb64 * a
        got = g["rgba"].reshape(-1, 4)
        err = np.abs(got[:, :3].astype(np.float64) - want).max()
        assert err <= PIXEL_TOL, f"frame {frames}: pixel error {err}"
        assert np.all(got[:, 3] == 1.0)
        assert np.abs(got - r["rgba"]).max() <= 2 * PIXEL_TOL
        imgs.append(g["rgba"].copy())
        prev = g["rgba"]
    # the jitter changes the rays: frame 0 differs from the unjittered frame
    u0 = uniforms([0, 0, 8], [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    plain = s.render(u0, W, H)["rgba"]
    assert not np.array_equal(plain, imgs[0]) and not np.array_equal(imgs[1], imgs[2])


def test_antialias_device_path_equals_host_path():
    _print("case antialias device path")
    raw, s, o = _scene()
    W, H = 64, 48
    dev = torch.device("cuda", 0)
    img = torch.zeros((H, W, 4), device=dev)
    prev = None
    for frames in range(3):
        u = uniforms([0.5, 0.2, 7.5], [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03,
                     accumulated_frames=frames)
        s.render_device(u, W, H, img.data_ptr(), antialiasing=True)
        host = s.render(u, W, H, antialiasing=True, rgba_prev=prev)["rgba"]
        torch.cuda.synchronize()
        np.testing.assert_array_equal(img.cpu().numpy().view(np.uint32), host.view(np.uint32))
        prev = host


def test_antialias_argument_errors():
    raw, s, o = _scene()
    u = uniforms([0, 0, 8], [0, 0, 0], 16, 16, k=6, accumulated_frames=0)
    with pytest.raises(gt.GTError):
        s.render(u, 16, 16, antialiasing=True, checkpoint_cap=8)
    dev = torch.device("cuda", 0)
    img = torch.zeros((16, 16, 4), device=dev)
    with pytest.raises(gt.GTError):
        gt.Batch(s, [u], [(16, 16)], [dict(rgba=img.data_ptr(), antialiasing=True)])
