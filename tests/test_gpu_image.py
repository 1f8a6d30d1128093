"""gt_tonemap on device buffers equals the host path byte for byte (the
curves share this library's FP64 exp/log), on a rendered frame and on
special values."""
import numpy as np
import pytest
import torch

from helpers import uniforms
import gtrace as gt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_device_tonemap_equals_host(mode):
    print(f"case device tonemap mode {mode}", flush=True)
    raw = gt.generate_raw(seed=0, config="c1")
    s = gt.Scene(raw, alpha_min=0.1)
    W, H = 193, 131
    u = uniforms([0, 0, 8], [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    frame = s.render(u, W, H)["rgba"]
    frame[0, :8, 0] = [np.nan, np.inf, -np.inf, -1.0, 0.0, 1.0, 2.0, 0.5 / 255.0]
    frame = frame * np.float32(1.7)  # some channels above 1
    d_in = torch.from_numpy(np.ascontiguousarray(frame)).to("cuda:0")
    d_out = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda:0")
    gt.tonemap_device(d_in.data_ptr(), W, H, d_out.data_ptr(), mode)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), gt.tonemap(frame, mode))


def test_render_into_reused_array():
    """Scene.render(out=...) writes the frame into the caller's array: equal to
    a render into a new array, across repeated calls."""
    print("case render out=", flush=True)
    raw = gt.generate_raw(seed=0, config="c1")
    s = gt.Scene(raw, alpha_min=0.1)
    W, H = 120, 90
    u = uniforms([0, 0, 8], [0, 0, 0], W, H, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    ref = s.render(u, W, H)["rgba"]
    buf = np.full((H, W, 4), 7.0, np.float32)
    for _ in range(2):
        got = s.render(u, W, H, out=buf)["rgba"]
        assert got is buf
        np.testing.assert_array_equal(buf.view(np.uint32), ref.view(np.uint32))
