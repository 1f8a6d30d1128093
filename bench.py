"""Benchmark of the forward Gaussian ray-tracing path on one MI355X.

Workload (one step = one forward frame, traceRays of the whole image):
  1,000,000 Gaussians, seeded 64-cluster mixture (centres uniform in a 20^3 box,
  offsets N(0, 0.8)), log-scales N(-4, 0.6), opacity logit N(0.5, 1.5), SH degree 3;
  one 1920x1080 pinhole camera, hFov 50 deg, at a seeded position inside the
  volume looking at the origin; alpha_min 0.01, k (hit_array_size) 16,
  T_min 0.001, t_min 1e-3, t_max 1e5, FP32.
Inputs are resident in HBM before the timed region; the output image is
written to device memory.  value = rays / time (Mrays/s).

roofline: algorithmic HBM bytes per frame
  B_fwd = 16 R + 44 H_eval + 192 H_comp
(R rays, H_eval evaluated k-buffer entries, H_comp composited entries: the
RGBA write, the 44-B mu/scale/rotation/opacity read per evaluated hit and
the 192-B SH read per composited hit), divided by the kernel's average
duration measured with HIP events on the launch stream.

fwd_bwd: forward with the backward checkpoint + dL/dRGBA of a squared-error
loss + the hand-written backward, on the same frame (and the recompute mode).
batched: 64 cameras of uneven resolution in one launch replayed from a
hipGraph, against the same cameras rendered one launch at a time.
fwd_bwd_c4: the same forward + backward on 200k single-cluster Gaussians, 800x800.
c5: 3M Gaussians, 16 views at 1280x720, forward with checkpoint + backward
per view with gradients accumulated across views, peak device memory (and
the recompute-mode step beside it).
device_build: the same 1M-Gaussian scene built on the GPU from device
parameters (rebuild time) and the forward frame time on that tree.
train_c4: one training step (forward with checkpoint, backward, Adam, scene
update) on the 200k-Gaussian 800x800 workload: refit of the host-built tree
(the default), and a GPU rebuild for comparison.

cpu_baseline: the oracle (the project's C restatement of the reference path,
BVH variant, all host threads up to 16) on a seeded 1/3 sample of the rays.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "webgpu-gaussiantracer_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_forward launch of the default frame (1M Gaussians,
# 1920x1080, k 16), from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
# this script scaled by the counters' reading for a 1 GiB device copy
# (x2.000 and x1.000): profiles/k_forward_1M_1080p_k16.md.  Reported as
# roofline.traffic for that workload only.
TRAFFIC_DEFAULT_FRAME = {"fetch_bytes": 543963750, "write_bytes": 171515637}
METRIC = "Mrays/s fwd and fwd+bwd, 1M Gaussians 1080p k=16; % of HBM roofline"


def camera_c2(seed=1):
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-10.0, 10.0, 3)
    return [float(x) for x in pos], [0.0, 0.0, 0.0]


BATCH_SIZES = [(128, 128), (320, 240), (640, 480), (1280, 720), (1920, 1080)]


def cameras_c3(n, seed=1):
    """n seeded poses on the sphere of radius 25 looking at the origin, with
    resolutions drawn uniformly from BATCH_SIZES."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    sizes = [BATCH_SIZES[i] for i in rng.integers(0, len(BATCH_SIZES), n)]
    return [list(25.0 * d) for d in dirs], sizes


def bench_batched(gt, scene, torch, dev, sp, K, n_cams, steps=3):
    """Batched forward: n_cams cameras of uneven resolution rendered by one
    launch replayed from a hipGraph, against the same cameras rendered one
    gt_render launch at a time (the plain single-camera kernel)."""
    poses, sizes = cameras_c3(n_cams)
    us = [gt.make_uniforms(gt.camera_look_at(p, [0.0, 0.0, 0.0]), w, h, hfov_deg=50.0, T_min=1e-3, alpha_min=0.01,
                           k=K) for p, (w, h) in zip(poses, sizes)]
    bufs = [torch.empty((h, w, 4), dtype=torch.float32, device=dev) for w, h in sizes]
    rays = sum(w * h for w, h in sizes)
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    saved = scene.get_tuning()
    scene.set_tuning()  # library default = the plain tile kernel

    def plain():
        for u, (w, h), b in zip(us, sizes, bufs):
            scene.render_device(u, w, h, b.data_ptr(), sp)
    ms_plain = timed(plain, steps)
    batch = gt.Batch(scene, us, sizes, [dict(rgba=b.data_ptr()) for b in bufs])
    tuned = gt.autotune_batch(scene, batch, use_graph=True, iters=1, stream_ptr=sp)
    ms_batch = timed(lambda: batch.render(sp, use_graph=True), steps)
    batch.close()
    scene.set_tuning(**{k: v for k, v in saved.items()})
    return {
        "workload": f"{n_cams} cameras on a sphere of radius 25, resolutions uniform from "
                    f"{BATCH_SIZES}, hFov 50, alpha_min 0.01, k {K}, T_min 0.001, one launch per frame "
                    f"replayed from a hipGraph",
        "rays": rays, "mrays": round(rays / ms_batch / 1e3, 3), "ms_per_frame": round(ms_batch, 3),
        "plain_mrays": round(rays / ms_plain / 1e3, 3), "plain_ms_per_frame": round(ms_plain, 3),
        "speedup_vs_plain": round(ms_plain / ms_batch, 3), "tuning": tuned["best"], "steps": steps,
    }


class MemSampler:
    """Peak device memory in use (hipMemGetInfo: total - free), sampled every ~1 ms on a thread."""

    def __init__(self, torch):
        import threading
        self.torch, self.peak, self._stop = torch, 0, threading.Event()
        self._t = threading.Thread(target=self._run, daemon=True)

    def _run(self):
        while not self._stop.is_set():
            free, total = self.torch.cuda.mem_get_info()
            self.peak = max(self.peak, total - free)
            self._stop.wait(0.001)

    def __enter__(self):
        self._t.start()
        return self

    def __exit__(self, *a):
        self._stop.set()
        self._t.join()
        free, total = self.torch.cuda.mem_get_info()
        self.peak = max(self.peak, total - free)


def measured_copy_gbs(torch, stream, nbytes=1 << 30, reps=5):
    """HBM bandwidth of a plain 1 GiB device-to-device copy on this box
    (bytes read + written per second), the measured counterpart of the
    nominal peak."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    ms = _events_ms(torch, stream, lambda: b.copy_(a), reps)
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def _events_ms(torch, stream, fn, n):
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def bench_fwd_bwd(gt, torch, dev, sp, raw, scene, u, W, H, steps, warmup, ms_fwd, st_fwd):
    """Forward + backward of loss = sum (render - target)^2, target rendered from
    the same scene with means perturbed by N(0, 0.01) (seed 3).  One step =
    forward with the per-ray checkpoint of composited ids, dL/dRGBA =
    2 (render - target), zeroing of the n x 59 gradient and gt_backward.  The
    recompute mode (no checkpoint; traversal re-run inside the backward over
    bands of a 256 MiB scratch) is timed beside it."""
    import numpy as np
    stream = torch.cuda.current_stream()
    R = W * H
    n = raw.shape[0]
    rng = np.random.Generator(np.random.PCG64(3))
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    target = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    st = gt.Scene(raw_t, alpha_min=0.01)
    st.render_device(u, W, H, target.data_ptr(), sp)
    torch.cuda.synchronize()
    st.close()
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    gimg = torch.empty_like(rgba)
    grads = torch.zeros((n, gt.GT_RAW_FLOATS), dtype=torch.float32, device=dev)
    ovf = torch.zeros(1, dtype=torch.int32, device=dev)
    ncomp = torch.empty(R, dtype=torch.int32, device=dev)
    scene.render_device(u, W, H, rgba.data_ptr(), sp, n_comp_ptr=ncomp.data_ptr(), comp_ids_ptr=None)
    torch.cuda.synchronize()
    most = int(ncomp.max().item())
    cap = 1 << max(0, most - 1).bit_length()
    cids = torch.empty((R, cap), dtype=torch.int32, device=dev)

    def fwd_ck():
        scene.render_device(u, W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(),
                            n_comp_ptr=ncomp.data_ptr(), comp_cap=cap)

    def loss_grad():
        torch.sub(rgba, target, out=gimg)
        gimg.mul_(2.0)
        grads.zero_()

    def bwd():
        gt.backward(scene, u, W, H, gimg.data_ptr(), grads.data_ptr(), cids.data_ptr(), ncomp.data_ptr(), cap,
                    ovf.data_ptr(), 0, sp)

    def step():
        fwd_ck()
        loss_grad()
        bwd()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    with MemSampler(torch) as mem:
        ms_fb = _events_ms(torch, stream, step, steps)
    ms_fwd_ck = _events_ms(torch, stream, fwd_ck, steps)
    ms_bwd = _events_ms(torch, stream, bwd, steps)
    step()  # one clean step for the reported loss and gradient norms
    torch.cuda.synchronize()
    assert int(ovf.item()) == 0
    loss = float(((rgba - target) ** 2).sum().item())
    gn = grads.norm(dim=0)

    def step_rc():
        scene.render_device(u, W, H, rgba.data_ptr(), sp)
        loss_grad()
        gt.backward(scene, u, W, H, gimg.data_ptr(), grads.data_ptr(), None, None, 0, ovf.data_ptr(), 0, sp)

    step_rc()
    torch.cuda.synchronize()
    rc_steps = max(1, steps // 4)
    with MemSampler(torch) as mem_rc:
        ms_rc = _events_ms(torch, stream, step_rc, rc_steps)
    h_comp, h_eval = st_fwd["h_comp"], st_fwd["h_eval"]
    B_fb = 16 * R + 44 * h_eval + 192 * h_comp + 16 * R + 44 * h_eval + (236 + 2 * 236) * h_comp
    return {
        "workload": f"loss = sum (render - target)^2, target from means + N(0, 0.01); forward with checkpoint, "
                    f"dL/dRGBA, backward; {n} Gaussians, {W}x{H}, k {u.k}",
        "mrays": round(R / ms_fb / 1e3, 3), "ms_per_step": round(ms_fb, 3),
        "ratio_vs_forward": round(ms_fb / ms_fwd, 3),
        "ms_forward_with_checkpoint": round(ms_fwd_ck, 3), "ms_backward": round(ms_bwd, 3),
        "checkpoint_cap": cap, "checkpoint_bytes": R * cap * 4 + R * 4, "max_composited_per_ray": most,
        "algorithmic_bytes": int(B_fb), "roofline_frac": round(B_fb / (ms_fb * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "peak_device_bytes": int(mem.peak), "loss": loss,
        "grad_norms": {"mean": float(gn[0:3].norm()), "sh": float(gn[3:51].norm()), "logit": float(gn[51]),
                       "log_scale": float(gn[52:55].norm()), "quat": float(gn[55:59].norm())},
        "recompute": {"ms_per_step": round(ms_rc, 3), "mrays": round(R / ms_rc / 1e3, 3),
                      "ratio_vs_forward": round(ms_rc / ms_fwd, 3), "peak_device_bytes": int(mem_rc.peak),
                      "steps": rc_steps},
        "steps": steps,
    }


def bench_c4(gt, torch, dev, sp, steps):
    """Forward + backward on 200,000 single-cluster Gaussians (N(0, 0.8)^3,
    log-scales N(-3.5, 0.5)), one 800x800 camera at (0, 0.5, 4) looking at
    the origin, hFov 50, alpha_min 0.01, k 16, T_min 0.001; same step and
    loss as the 1M-Gaussian forward + backward."""
    raw = gt.generate_raw(seed=0, config="c4")
    scene = gt.Scene(raw, alpha_min=0.01)
    W = H = 800
    u = gt.make_uniforms(gt.camera_look_at([0.0, 0.5, 4.0], [0.0, 0.0, 0.0]), W, H, hfov_deg=50.0, T_min=1e-3,
                         alpha_min=0.01, k=16)
    stream = torch.cuda.current_stream()
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    stats = torch.zeros(8, dtype=torch.int64, device=dev)
    scene.render_device(u, W, H, rgba.data_ptr(), sp, stats.data_ptr())
    torch.cuda.synchronize()
    st = dict(zip(gt.STAT_NAMES, [int(x) for x in stats.cpu().tolist()]))
    ms_f = _events_ms(torch, stream, lambda: scene.render_device(u, W, H, rgba.data_ptr(), sp), steps)
    r = bench_fwd_bwd(gt, torch, dev, sp, raw, scene, u, W, H, steps, 1, ms_f, st)
    r["ms_forward"] = round(ms_f, 3)
    r["workload"] = ("200000 Gaussians single cluster, 800x800, k 16: " + r["workload"].split("; ")[0])
    return r


def bench_c5(gt, torch, dev, sp, steps):
    """Batched forward + backward at scale: 3M Gaussians (256-cluster mixture in
    a 40^3 box, log-scales N(-4.5, 0.7)), 16 cameras at 1280x720 on a sphere of
    radius 25 (seed 1); per view: forward with the composited-id checkpoint
    (one 1280x720 x 256 buffer reused by the views), dL/dRGBA of
    sum (render - target)^2 (target from means + N(0, 0.01), seed 3), and the
    backward accumulating into one 3M x 59 gradient with FP32 atomics.  The
    recompute-mode step (no checkpoint; the traversal re-runs in bands) is
    timed beside it, and replaces it if a ray composites more than 256 hits.
    Reports the peak device memory in use during the timed steps."""
    import numpy as np
    stream = torch.cuda.current_stream()
    raw = gt.generate_raw(seed=0, config="c5")
    n = raw.shape[0]
    scene = gt.Scene(raw, alpha_min=0.01)
    rng = np.random.Generator(np.random.PCG64(3))
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0.0, 0.01, (n, 3)).astype(np.float32)
    W, H, V = 1280, 720, 16
    rngc = np.random.Generator(np.random.PCG64(1))
    dirs = rngc.normal(size=(V, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    us = [gt.make_uniforms(gt.camera_look_at(list(25.0 * d), [0.0, 0.0, 0.0]), W, H, hfov_deg=50.0, T_min=1e-3,
                           alpha_min=0.01, k=16) for d in dirs]
    targets = torch.empty((V, H, W, 4), dtype=torch.float32, device=dev)
    ts = gt.Scene(raw_t, alpha_min=0.01)
    for i, u in enumerate(us):
        ts.render_device(u, W, H, targets[i].data_ptr(), sp)
    torch.cuda.synchronize()
    ts.close()
    del raw_t
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    gimg = torch.empty_like(rgba)
    grads = torch.zeros((n, gt.GT_RAW_FLOATS), dtype=torch.float32, device=dev)
    ovf = torch.zeros(1, dtype=torch.int32, device=dev)
    cap = 256
    cids = torch.empty((H * W, cap), dtype=torch.int32, device=dev)
    ncomp = torch.empty(H * W, dtype=torch.int32, device=dev)
    max_comp = torch.zeros((), dtype=torch.int32, device=dev)

    def fwd():
        for u in us:
            scene.render_device(u, W, H, rgba.data_ptr(), sp)

    def step_rc():
        grads.zero_()
        for i, u in enumerate(us):
            scene.render_device(u, W, H, rgba.data_ptr(), sp)
            torch.sub(rgba, targets[i], out=gimg)
            gimg.mul_(2.0)
            gt.backward(scene, u, W, H, gimg.data_ptr(), grads.data_ptr(), None, None, 0, ovf.data_ptr(), 0, sp)

    def step_ck(track_max=False):
        grads.zero_()
        for i, u in enumerate(us):
            scene.render_device(u, W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(),
                                n_comp_ptr=ncomp.data_ptr(), comp_cap=cap)
            if track_max:
                torch.maximum(max_comp, ncomp.max(), out=max_comp)
            torch.sub(rgba, targets[i], out=gimg)
            gimg.mul_(2.0)
            gt.backward(scene, u, W, H, gimg.data_ptr(), grads.data_ptr(), cids.data_ptr(), ncomp.data_ptr(), cap,
                        ovf.data_ptr(), 0, sp)

    rays = V * W * H
    step_rc()
    torch.cuda.synchronize()
    assert int(ovf.item()) == 0
    with MemSampler(torch) as mem_rc:
        ms_rc = _events_ms(torch, stream, step_rc, max(1, steps // 2))
    norm_rc = float(grads.norm().item())
    step_ck(track_max=True)
    torch.cuda.synchronize()
    checkpoint = int(ovf.item()) == 0
    if checkpoint:
        with MemSampler(torch) as mem:
            ms = _events_ms(torch, stream, step_ck, steps)
        mode = f"checkpoint (cap {cap}, one buffer reused by the views)"
    else:
        ms, mem = ms_rc, mem_rc
        mode = f"recompute (a ray composited more than {cap} hits)"
    assert int(ovf.item()) == 0 or not checkpoint
    ms_f = _events_ms(torch, stream, fwd, steps)
    info = scene.info()
    return {
        "workload": f"{n} Gaussians 256-cluster mixture, {V} cameras {W}x{H} on a sphere of radius 25, alpha_min "
                    f"0.01, k 16, T_min 0.001; per view forward + dL/dRGBA + backward, gradients accumulated "
                    f"across views",
        "mode": mode, "max_composited_per_ray": int(max_comp.item()),
        "rays_per_step": rays, "ms_per_step": round(ms, 3), "mrays": round(rays / ms / 1e3, 3),
        "ms_forward_only": round(ms_f, 3), "ratio_vs_forward": round(ms / ms_f, 3),
        "peak_device_bytes": int(mem.peak), "scene_device_bytes": info["device_bytes"],
        "checkpoint_bytes": int(cids.numel() * 4 + ncomp.numel() * 4),
        "gradient_bytes": n * gt.GT_RAW_FLOATS * 4, "grad_norm": float(grads.norm().item()),
        "recompute": {"ms_per_step": round(ms_rc, 3), "mrays": round(rays / ms_rc / 1e3, 3),
                      "ratio_vs_forward": round(ms_rc / ms_f, 3), "peak_device_bytes": int(mem_rc.peak),
                      "grad_norm": norm_rc, "steps": max(1, steps // 2)},
        "steps": steps,
    }


def bench_device_build(gt, torch, dev, sp, raw, u, W, H, steps):
    """The same scene built on the GPU from device-resident parameters
    (gt_scene_create_device: activations, proxies, Morton radix-tree BVH),
    rebuild time (gt_scene_update_device) and the forward frame time on the
    GPU-built tree."""
    stream = torch.cuda.current_stream()
    x = torch.from_numpy(raw).to(dev)
    torch.cuda.synchronize()
    scene = gt.Scene(raw_dev=x.data_ptr(), n=raw.shape[0], alpha_min=0.01, stream_ptr=sp)
    reb = []
    for _ in range(3):
        t0 = time.perf_counter()
        scene.update_device(x.data_ptr(), sp)
        reb.append((time.perf_counter() - t0) * 1e3)
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    for _ in range(2):
        scene.render_device(u, W, H, rgba.data_ptr(), sp)
    ms = _events_ms(torch, stream, lambda: scene.render_device(u, W, H, rgba.data_ptr(), sp), steps)
    info = scene.info()
    scene.close()
    return {"rebuild_ms": round(sorted(reb)[1], 3), "forward_ms": round(ms, 4),
            "mrays": round(W * H / ms / 1e3, 3), "n_nodes": info["n_nodes"], "bvh_depth": info["bvh_depth"],
            "max_stack": info["max_stack"], "device_bytes": info["device_bytes"]}


def bench_train_c4(gt, torch, dev, sp, steps):
    """One training step at 200k single-cluster Gaussians, one 800x800 view
    (the forward + backward workload above): forward with checkpoint,
    dL/dRGBA, backward, Adam with the 3DGS learning rates, then the scene
    update.  Measured twice: the default trainer (host binned-SAH tree built
    once, refit after each step) and the rebuild trainer (GPU Morton tree
    rebuilt after each step).  Both give the same renders and gradients."""
    import numpy as np
    from gtrace.train import Trainer
    raw = gt.generate_raw(seed=0, config="c4")
    rng = np.random.Generator(np.random.PCG64(3))
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0.0, 0.01, (raw.shape[0], 3)).astype(np.float32)
    W = H = 800
    u = gt.make_uniforms(gt.camera_look_at([0.0, 0.5, 4.0], [0.0, 0.0, 0.0]), W, H, hfov_deg=50.0, T_min=1e-3,
                         alpha_min=0.01, k=16)
    ts = gt.Scene(raw_t, alpha_min=0.01)
    target = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    ts.render_device(u, W, H, target.data_ptr(), sp)
    torch.cuda.synchronize()
    ts.close()

    def run(build, refit, nsteps):
        x = torch.from_numpy(raw).to(dev)
        tr = Trainer(x, [(u, W, H, target)], alpha_min=0.01, build=build, refit=refit)
        loss0 = tr.step()
        tr.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = [tr.step() for _ in range(nsteps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / nsteps * 1e3
        r = {"ms_per_step": round(ms, 3), "mrays": round(W * H / ms / 1e3, 3),
             "update_ms": round(tr.scene.info()["build_ms"], 3), "rebuilds": tr.rebuilds,
             "loss_first": round(loss0, 3), "loss_last": round(losses[-1], 3), "steps": nsteps}
        tr.close()
        return r

    out = {"workload": "200000 Gaussians single cluster, one 800x800 view, k 16; step = forward with checkpoint + "
                       "dL/dRGBA + backward + Adam (3DGS learning rates) + scene update (host SAH tree built once, "
                       "refit each step; fallback GPU rebuild when the proxy set changes)"}
    out.update(run("host", True, steps))
    rb = run("gpu", False, max(3, steps // 2))
    rb["workload"] = "the same step with a GPU rebuild (radix tree with PLOC treelets) instead of the refit"
    out["gpu_rebuild"] = rb
    out["note"] = ("update_ms is the wall time of the last update call, including the wait for the queued "
                   "backward at its first host read")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-camera measurement")
    ap.add_argument("--batch-cams", type=int, default=64)
    ap.add_argument("--no-backward", action="store_true", help="skip the forward+backward measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the 3M-Gaussian 16-view forward+backward")
    ap.add_argument("--no-c4", action="store_true", help="skip the 200k-Gaussian 800x800 forward+backward")
    ap.add_argument("--no-train", action="store_true", help="skip the device build and training-step measurements")
    ap.add_argument("--cpu-sample", type=int, default=3, help="time 1/N of the rays on the CPU")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": a.numel() * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(rays would shard over GPUs by tile ranges with no exchange, each GPU holding the whole scene)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import gtrace as gt

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    W, H, K = args.width, args.height, args.k
    R = W * H

    t0 = time.time()
    raw = gt.generate_raw(n=args.n, seed=0, config="c2")
    t_gen = time.time() - t0
    t0 = time.time()
    scene = gt.Scene(raw, alpha_min=0.01)
    t_build = time.time() - t0
    info = scene.info()
    pos, tgt = camera_c2()
    tw = gt.camera_look_at(pos, tgt)
    u = gt.make_uniforms(tw, W, H, hfov_deg=50.0, t_min=1e-3, t_max=1e5, T_min=1e-3, alpha_min=0.01, k=K,
                         sh_degree_max=3)

    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    stats = torch.zeros(8, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    # counters for the algorithmic byte count (separate, untimed launch)
    scene.render_device(u, W, H, rgba.data_ptr(), sp, stats.data_ptr())
    torch.cuda.synchronize()
    st = dict(zip(gt.STAT_NAMES, [int(x) for x in stats.cpu().tolist()]))

    for _ in range(args.warmup):
        scene.render_device(u, W, H, rgba.data_ptr(), sp)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        scene.render_device(u, W, H, rgba.data_ptr(), sp)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    mrays = R / (ms_step * 1e-3) / 1e6

    B_fwd = 16 * R + 44 * st["h_eval"] + 192 * st["h_comp"]
    achieved = B_fwd / (ms_step * 1e-3) / 1e9
    img = rgba.cpu().numpy()
    copy_gbs = measured_copy_gbs(torch, stream)
    out = {
        "metric": METRIC,
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, scene seed 0, camera seed 1)",
        "config": {
            "workload": f"forward render, {args.n} Gaussians 64-cluster mixture SH3, {W}x{H}, hFov 50, "
                        f"alpha_min 0.01, k {K}, T_min 0.001",
            "gaussians": args.n, "width": W, "height": H, "k": K, "camera_pos": [round(x, 4) for x in pos],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": (TRAFFIC_DEFAULT_FRAME["fetch_bytes"] + TRAFFIC_DEFAULT_FRAME["write_bytes"]
                        if (args.n, W, H, K) == (1000000, 1920, 1080, 16) else None),
            "algorithmic_bytes": int(B_fwd), "kernel": "k_forward<16, 0, false, false, false, true> (packet traversal)",
            "measured_copy_gbs": round(copy_gbs, 1), "frac_of_measured_copy": round(achieved / copy_gbs, 5),
        },
        "stats": st,
        "wall_ms_per_step": round(wall / args.steps * 1e3, 4),
        "scene": {"n_proxies": info["n_proxies"], "n_nodes": info["n_nodes"], "bvh_depth": info["bvh_depth"],
                  "max_stack": info["max_stack"], "device_bytes": info["device_bytes"],
                  "host_build_s": round(t_build, 3), "generate_s": round(t_gen, 3)},
        "image_mean": float(img[..., :3].mean()),
    }

    # host-buffer path (what the Node addon uses): device temporaries, render,
    # copy-back over PCIe, synchronous -- reported beside value, never as it
    import numpy as np
    hp, hr = [], []
    frame = np.zeros((H, W, 4), np.float32)
    for _ in range(3):
        h0 = time.perf_counter()
        scene.render(u, W, H)
        hp.append(time.perf_counter() - h0)
        h0 = time.perf_counter()
        scene.render(u, W, H, out=frame)
        hr.append(time.perf_counter() - h0)
    hp_s, hr_s = sorted(hp)[1], sorted(hr)[1]
    out["host_path"] = {"ms_per_frame": round(hr_s * 1e3, 3), "mrays": round(R / hr_s / 1e6, 3),
                        "ms_per_frame_new_array": round(hp_s * 1e3, 3),
                        "note": "gt_render with host buffers: render, then PCIe copy-back into a host array reused "
                                "across frames (new_array: a fresh array per frame, whose pages fault in during the "
                                "copy), median of 3"}

    if not args.no_backward:
        out["fwd_bwd"] = bench_fwd_bwd(gt, torch, dev, sp, raw, scene, u, W, H, max(2, args.steps // 2),
                                       min(args.warmup, 2), ms_step, st)
    if not args.no_train:
        out["device_build"] = bench_device_build(gt, torch, dev, sp, raw, u, W, H, max(2, args.steps // 2))
        out["device_build"]["host_build_ms"] = round(t_build * 1e3, 1)
    if not args.no_batched:
        out["batched"] = bench_batched(gt, scene, torch, dev, sp, K, args.batch_cams)
    if not args.no_backward and not args.no_c4:
        out["fwd_bwd_c4"] = bench_c4(gt, torch, dev, sp, 5)
    if not args.no_backward and not args.no_train:
        out["train_c4"] = bench_train_c4(gt, torch, dev, sp, 10)
    if not args.no_c5:
        out["c5"] = bench_c5(gt, torch, dev, sp, 2)

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        o = orc.OracleScene(raw, 0.01)
        o.prebuild_bvh()
        p = orc.params_from_uniforms(u, W, H)
        rng = np.random.Generator(np.random.PCG64(2))
        n_s = max(1, R // args.cpu_sample)  # a seeded 1/N sample of the frame
        rays = np.sort(rng.choice(R, n_s, replace=False)).astype(np.uint32)
        c0 = time.perf_counter()
        rr = orc.render(o, p, threads=threads, ray_list=rays)
        cs = time.perf_counter() - c0
        out["cpu_baseline"] = {
            "value": round(n_s / cs / 1e6, 5), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n_s} seeded random rays (1/{args.cpu_sample} of the frame) of the same workload, "
                      f"oracle BVH variant (CPU restatement of the reference path, not the reference binary), "
                      f"{cs:.2f} s",
            "h_eval_per_ray": rr["stats"]["h_eval"] / n_s,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
