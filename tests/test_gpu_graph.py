"""A one-camera render (device buffers) captured into a graph and replayed.

The frame's proxy-rectangle table and its per-Gaussian colour buffer are
allocated per stream on its first render outside a capture.  Captured on a
stream that has not rendered yet, the render takes neither (nothing may be
allocated during a capture: each composited hit evaluates its SH record);
captured on a stream that has, the rectangle and colour passes are part of
the graph.
Either way every replay equals the eager render bit for bit.  The replays
run with eager work on the same and another stream between them, and the
output is cleared by a kernel before each one.
"""
import numpy as np
import pytest
import torch

from helpers import uniforms
import gtrace as gt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("warm", [False, True])
def test_single_camera_render_replays_from_a_graph(warm):
    print(f"case graph replay warm={warm}", flush=True)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    raw = gt.generate_raw(n=60000, seed=3, config="c1")
    s = gt.Scene(raw, alpha_min=0.05)
    s.set_tuning(rects=1, colors=1)  # the culling and the colour pass on for this small frame
    W, H = 160, 120
    u = uniforms([0.3, -0.2, 7.0], [0, 0, 0], W, H, hfov_deg=55, alpha_min=0.05, k=12, T_min=0.01)
    u_other = uniforms([-0.5, 0.4, 6.0], [0, 0, 0], W, H, hfov_deg=55, alpha_min=0.05, k=12, T_min=0.01)

    side = torch.cuda.Stream()
    ref = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
    with torch.cuda.stream(side):
        s.render_device(u, W, H, ref.data_ptr(), side.cuda_stream)
    torch.cuda.synchronize()
    host = s.render(u, W, H)["rgba"]  # the host-buffer path on the default stream
    assert np.array_equal(ref.cpu().numpy().view(np.uint32), host.view(np.uint32))

    cap = torch.cuda.Stream()
    out = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
    if warm:  # this stream's rectangle table exists before the capture
        with torch.cuda.stream(cap):
            s.render_device(u, W, H, out.data_ptr(), cap.cuda_stream)
        torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cap):
        s.render_device(u, W, H, out.data_ptr(), cap.cuda_stream)
    torch.cuda.synchronize()

    other = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), ref.view(torch.int32)), f"replay {rep}"
        # eager renders between the replays: another frame on the capture
        # stream and on a side stream
        with torch.cuda.stream(cap):
            s.render_device(u_other, W, H, other.data_ptr(), cap.cuda_stream)
        with torch.cuda.stream(side):
            s.render_device(u_other, W, H, other.data_ptr(), side.cuda_stream)
        torch.cuda.synchronize()
    del g
    s.close()
