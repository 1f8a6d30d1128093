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
  B_fwd = 16 R + 44 H_eval + 16 H_comp + 224 N
(R rays, H_eval evaluated k-buffer entries, H_comp composited entries, N
Gaussians: the RGBA write, the 44-B mu/scale/rotation/opacity read per
evaluated hit, the 16-B colour read per composited hit, and the colour pass
that reads each Gaussian's 16-B mu and 192-B SH record and writes its 16-B
colour; a frame with fewer than N / 2 rays reads the 192-B SH record per
composited hit instead, 16 R + 44 H_eval + 192 H_comp), divided by the
frame's average duration measured with HIP events on the launch stream.  roofline.traffic is
the kernel's FETCH_SIZE + WRITE_SIZE per launch, measured in the same run by
rocprofv3 --pmc child processes and scaled by what the counters read for a
plain 1 GiB device copy (null, with the reason, when rocprofv3 is missing or
fails).

fwd_bwd: forward with the backward checkpoint + dL/dRGBA of a squared-error
loss + the hand-written backward, on the same frame (and the recompute mode),
with a roofline per kernel group from the bytes those kernels must move.
batched: 64 cameras of uneven resolution in one launch replayed from a
hipGraph, for alpha_min 0.01 and 0.1 and k 4, 8, 16, 32, each with the ray
tile shape the autotuner picks and checked bit for bit against the same
cameras rendered one launch at a time (the plain kernel at that k).
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
BVH variant) on every core the process may use, on a seeded 1/3 sample of
the rays.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "webgpu-gaussiantracer_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s fwd and fwd+bwd, 1M Gaussians 1080p k=16; % of HBM roofline"
FWD_KERNEL = ("k_forward<32, 3, false, false, false, true, true> (packet traversal, persistent kernel, 32-slot buffer "
              "serving rounds of k, per-Gaussian colours precomputed by k_colors)")


def forward_bytes(R, h_eval, h_comp, n):
    """Algorithmic bytes of one frame: the composite reads a 16-B colour per
    hit after a colour pass over the N Gaussians (224 B each) when the frame
    has at least N / 2 rays (the library's default), else a 192-B SH record
    per hit."""
    if 2 * R >= n:
        return 16 * R + 44 * h_eval + 16 * h_comp + 224 * n
    return 16 * R + 44 * h_eval + 192 * h_comp


def _counter_csv_rows(outdir):
    """Rows of every rocprofv3 counter_collection.csv under outdir."""
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _per_dispatch(rows, counter):
    """[(kernel name, value)] per dispatch in dispatch order, summed over the
    counter's rows of that dispatch."""
    acc = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        key = int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0)
        name, v = acc.get(key, (r.get("Kernel_Name", ""), 0.0))
        acc[key] = (name, v + float(r.get("Counter_Value") or 0.0))
    return [acc[k] for k in sorted(acc)]


def measure_traffic(args, timeout_s=120):
    """HBM bytes of the forward kernel (first two plain launches of the
    default frame) and of the backward's kernels (k_colors, k_backward,
    k_finalize of one checkpoint-mode call), from rocprofv3 --pmc FETCH_SIZE
    and WRITE_SIZE child runs of this script (--traffic-probe), each counter
    scaled by its reading for a plain 1 GiB device copy (--copy-calib).  Run
    before this process touches the GPU; returns (dict, None) or (None,
    reason)."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                          else None)
    if prof is None:
        return None, "rocprofv3 not found"
    base = [sys.executable, os.path.abspath(__file__)]
    probe = base + ["--traffic-probe", "--n", str(args.n), "--width", str(args.width), "--height",
                    str(args.height), "--k", str(args.k)]
    calib = base + ["--copy-calib"]
    out = {}
    tmp = tempfile.mkdtemp(prefix="gt_traffic_")
    env = dict(os.environ, TMPDIR=tempfile.gettempdir())
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            for name, cmd in (("probe", probe), ("calib", calib)):
                d = os.path.join(tmp, f"{name}_{counter}")
                r = subprocess.run([prof, "--pmc", counter, "-d", d, "-o", "run", "-f", "csv", "--"] + cmd,
                                   capture_output=True, text=True, timeout=timeout_s, env=env)
                if r.returncode != 0:
                    return None, f"rocprofv3 --pmc {counter} ({name}) exited {r.returncode}: {r.stderr[-300:]}"
                out[(name, counter)] = _per_dispatch(_counter_csv_rows(d), counter)
        # VALU issue of the same launches: instructions, active quad-cycles,
        # thread-cycles and the busy cycles of the XCDs (one pass; SQ and
        # GRBM blocks within their counter limits)
        d = os.path.join(tmp, "probe_valu")
        r = subprocess.run([prof, "--pmc"] + list(VALU_COUNTERS) + ["-d", d, "-o", "run", "-f", "csv", "--"] + probe,
                           capture_output=True, text=True, timeout=timeout_s, env=env)
        if r.returncode != 0:
            return None, f"rocprofv3 --pmc VALU counters exited {r.returncode}: {r.stderr[-300:]}"
        rows = _counter_csv_rows(d)
        out["valu"] = {c: _per_dispatch(rows, c) for c in VALU_COUNTERS}
        out["agent"] = _agent_info(d)
        d = os.path.join(tmp, "probe_trace")
        r = subprocess.run([prof, "--kernel-trace", "-d", d, "-o", "run", "-f", "csv", "--"] + probe,
                           capture_output=True, text=True, timeout=timeout_s, env=env)
        if r.returncode != 0:
            return None, f"rocprofv3 --kernel-trace (probe) exited {r.returncode}: {r.stderr[-300:]}"
        out["trace"] = _kernel_trace_ms(d)
    except subprocess.TimeoutExpired:
        return None, f"rocprofv3 pass took more than {timeout_s} s"
    except OSError as e:
        return None, f"rocprofv3 could not run: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    res = {"copy_bytes": 1 << 30}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = [v for n, v in out[("calib", counter)] if "copy" in n.lower()] or \
            [max((v for _, v in out[("calib", counter)]), default=0.0)]
        kib = sum(cal) / len(cal)
        if kib <= 0:
            return None, f"no {counter} reading for the calibration copy"
        scale = (1 << 30) / (kib * 1024.0)  # the copy reads and writes 1 GiB
        fwd = [v for n, v in out[("probe", counter)] if "k_forward<" in n][:2]
        if len(fwd) < 2:
            return None, f"no {counter} reading for k_forward"
        res[counter] = {"scale": round(scale, 4), "calib_kib": round(kib, 1),
                        "k_forward_bytes": int(sum(fwd) / len(fwd) * 1024.0 * scale)}
        bwd = {}
        for kname in ("k_colors", "k_backward", "k_finalize"):
            v = [x for n, x in out[("probe", counter)] if kname in n]
            bwd[kname] = int(sum(v) * 1024.0 * scale)
        res[counter]["backward_bytes"] = bwd
    res["kernel_ms"] = out["trace"]
    res["valu"] = _valu_issue(out["valu"], out["agent"], out["trace"].get("k_forward", [])[:2])
    return res, None


VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "GRBM_GUI_ACTIVE")


def _agent_info(outdir):
    """(SIMDs, XCDs) of the GPU agent from rocprofv3's agent_info.csv."""
    import csv
    import glob
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*agent_info.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if int(r.get("Simd_Count") or 0) > 0:
                    return int(r["Simd_Count"]), int(r.get("Num_Xcc") or 1)
    return None


def _valu_issue(valu, agent, trace_ms):
    """VALU issue of the forward kernel (mean of the first two plain launches):
    instructions, the fraction of the launch every SIMD's VALU is busy
    (SQ_ACTIVE_INST_VALU counts quad-cycles: a wave64 VALU instruction
    occupies a SIMD for 4 cycles), lane efficiency and the clock over the
    launch (GRBM_GUI_ACTIVE summed over the XCDs, per XCD, over the traced
    duration of the same launches)."""
    if not agent:
        return None
    simds, xcds = agent
    v = {}
    for c in VALU_COUNTERS:
        fwd = [x for n, x in valu[c] if "k_forward<" in n][:2]
        if len(fwd) < 2:
            return None
        v[c] = sum(fwd) / len(fwd)
    cycles = v["GRBM_GUI_ACTIVE"] / xcds
    out = {"valu_insts": int(v["SQ_INSTS_VALU"]),
           "busy_frac": round(4.0 * v["SQ_ACTIVE_INST_VALU"] / (simds * cycles), 4) if cycles else None,
           "lane_efficiency": round(v["SQ_THREAD_CYCLES_VALU"] / (64.0 * v["SQ_ACTIVE_INST_VALU"]), 4),
           "cycles_per_xcd": int(cycles), "simds": simds, "xcds": xcds}
    if len(trace_ms) >= 2:
        out["clock_ghz"] = round(cycles / (sum(trace_ms) / len(trace_ms) * 1e-3) / 1e9, 3)
    return out


def _kernel_trace_ms(outdir):
    """{kernel: [duration ms per dispatch]} from rocprofv3 kernel_trace.csv files."""
    import csv
    import glob
    ms = {}
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*kernel_trace.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                for key in ("k_forward<", "k_colors", "k_backward", "k_finalize", "k_zero_f32"):
                    if key in name:
                        d = (float(r["End_Timestamp"]) - float(r["Start_Timestamp"])) * 1e-6
                        ms.setdefault(key.rstrip("<"), []).append(round(d, 4))
    return ms


def traffic_probe(args):
    """--traffic-probe: two plain frames of the default workload, then one
    forward with the backward checkpoint and one backward (what
    measure_traffic profiles)."""
    import torch
    import gtrace as gt
    dev = torch.device("cuda:0")
    W, H, K = args.width, args.height, args.k
    raw = gt.generate_raw(n=args.n, seed=0, config="c2")
    scene = gt.Scene(raw, alpha_min=0.01)
    pos, tgt = camera_c2()
    u = gt.make_uniforms(gt.camera_look_at(pos, tgt), W, H, hfov_deg=50.0, t_min=1e-3, t_max=1e5, T_min=1e-3,
                         alpha_min=0.01, k=K, sh_degree_max=3)
    sp = torch.cuda.current_stream().cuda_stream
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    for _ in range(2):
        scene.render_device(u, W, H, rgba.data_ptr(), sp)
    R = W * H
    ncomp = torch.empty(R, dtype=torch.int32, device=dev)
    scene.render_device(u, W, H, rgba.data_ptr(), sp, n_comp_ptr=ncomp.data_ptr())
    torch.cuda.synchronize()
    cap = 1 << max(0, int(ncomp.max().item()) - 1).bit_length()
    cids = torch.empty((R, cap), dtype=torch.int32, device=dev)
    scene.render_device(u, W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(), n_comp_ptr=ncomp.data_ptr(),
                        comp_cap=cap)
    g = torch.ones_like(rgba)
    grads = torch.zeros((raw.shape[0], gt.GT_RAW_FLOATS), dtype=torch.float32, device=dev)
    gt.backward(scene, u, W, H, g.data_ptr(), grads.data_ptr(), cids.data_ptr(), ncomp.data_ptr(), cap, None, 0, sp)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "done"}), flush=True)


def cpu_info():
    """(threads the process may run on, nproc, CPU model, cgroup CPU quota)."""
    import math
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = aff if quota is None else max(1, min(aff, int(math.floor(quota))))
    return threads, os.cpu_count(), model, quota, aff


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


def bench_batched(gt, raw, scene, torch, dev, sp, n_cams, alphas=(0.01, 0.1), ks=(4, 8, 16, 32)):
    """Batched forward: n_cams cameras of uneven resolution rendered by one
    launch replayed from a hipGraph, for alpha_min 0.01 and 0.1 and
    hit_array_size 4, 8, 16, 32; for each, the ray-tile shape (4, 8, 16 or
    32 pixels wide) is picked by gtrace.autotune_batch_k, and the batch is
    checked bit for bit against the same cameras rendered one gt_render
    launch at a time at that k (the plain kernel), whose time is reported
    beside it.  The headline entry is alpha_min 0.01, k 16."""
    poses, sizes = cameras_c3(n_cams)
    rays = sum(w * h for w, h in sizes)
    sweep, fastest = [], {}
    for alpha in alphas:
        sc = scene if alpha == 0.01 else gt.Scene(raw, alpha_min=alpha)
        us = [gt.make_uniforms(gt.camera_look_at(p, [0.0, 0.0, 0.0]), w, h, hfov_deg=50.0, T_min=1e-3,
                               alpha_min=alpha, k=16) for p, (w, h) in zip(poses, sizes)]
        res, fk = gt.autotune_batch_k(sc, us, sizes, ks=ks, tile_ws=(4, 8, 16, 32), iters=2, stream_ptr=sp)
        fastest[str(alpha)] = fk
        for r in res:
            sweep.append({"alpha_min": alpha, "k": r["k"], "mrays": round(rays / r["ms"] / 1e3, 3),
                          "ms_per_frame": round(r["ms"], 3), "tile_w": r["tile_w"],
                          "plain_mrays": round(rays / r["plain_ms"] / 1e3, 3),
                          "plain_ms_per_frame": round(r["plain_ms"], 3),
                          "speedup_vs_plain": round(r["plain_ms"] / r["ms"], 3),
                          "equal_to_plain": r["equal_to_plain"]})
        if sc is not scene:
            sc.close()
    head = ([r for r in sweep if r["alpha_min"] == 0.01 and r["k"] == 16] or sweep)[0]
    return {
        "workload": f"{n_cams} cameras on a sphere of radius 25, resolutions uniform from {BATCH_SIZES}, hFov 50, "
                    f"T_min 0.001; one launch per frame replayed from a hipGraph (persistent packet kernel, ray "
                    f"tile shape chosen by the autotuner from widths 4, 8, 16, 32); headline alpha_min 0.01, k 16",
        "rays": rays, "mrays": head["mrays"], "ms_per_frame": head["ms_per_frame"],
        "plain_mrays": head["plain_mrays"], "plain_ms_per_frame": head["plain_ms_per_frame"],
        "speedup_vs_plain": head["speedup_vs_plain"], "tuning": {"path": 4, "tile_w": head["tile_w"]},
        "sweep": sweep, "fastest_k": fastest,
        "note": "plain = the same cameras as one gt_render launch each at the same k; every batch equals its plain "
                "renders bit for bit; k changes the image, so the fastest k is reported, not imposed",
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


def backward_counts(torch, cids, ncomp, W, H, n):
    """From a frame's checkpoint: composited hits H_comp, rays with hits
    R_hit, distinct composited Gaussians N_touched, and distinct (16x16
    pixel block, Gaussian) pairs P (the backward flushes its per-block
    accumulators once per pair)."""
    R = W * H
    cap = cids.shape[1]
    nc = ncomp.clamp(max=cap)
    valid = torch.arange(cap, device=cids.device)[None, :] < nc[:, None]
    pix = torch.arange(R, device=cids.device, dtype=torch.int64)[:, None].expand(R, cap)[valid]
    gid = cids[valid].to(torch.int64)
    bx = (W + 15) // 16
    block = (pix // W // 16) * bx + (pix % W) // 16
    out = {"H_comp": int(gid.numel()), "R_hit": int((ncomp > 0).sum().item()),
           "N_touched": int(torch.unique(gid).numel()), "P": int(torch.unique(block * n + gid).numel())}
    del pix, gid, block, valid
    return out


def backward_bytes(c, R, n, W, H):
    """Algorithmic bytes of one checkpoint-mode gt_backward call by kernel."""
    nb = ((W + 15) // 16) * ((H + 15) // 16)  # 16x16-pixel blocks
    b = {"k_zero_f32": 64 * n, "k_colors": 224 * n,
         "k_bw_cost": 4 * R + 4 * nb, "k_bw_order": 12 * nb,  # counts; block costs; order
         "k_backward": 4 * R + 16 * c["R_hit"] + 136 * c["H_comp"] + 112 * c["P"],
         "k_finalize": 64 * n + 712 * c["N_touched"]}
    b["total"] = sum(b.values())
    return b


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
    cnt = backward_counts(torch, cids, ncomp, W, H, n)
    B = backward_bytes(cnt, R, n, W, H)
    B_fwd_ck = forward_bytes(R, h_eval, h_comp, n) + 4 * h_comp + 4 * R  # + the checkpoint's ids and counts
    B_loss = 80 * R + 236 * n  # dL/dRGBA = 2 (render - target) and zeroing the N x 59 gradient
    B_step = B_fwd_ck + B_loss + B["total"]

    def roof(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"algorithmic_bytes": int(nbytes), "ms": round(ms, 4), "achieved_gbs": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 5)}
    roofline = {
        "forward_with_checkpoint": roof(B_fwd_ck, ms_fwd_ck),
        "backward": dict(roof(B["total"], ms_bwd), per_kernel_bytes={k: int(v) for k, v in B.items() if k != "total"}),
        "step": dict(roof(B_step, ms_fb), note="step-level: forward with checkpoint + dL/dRGBA + gradient zeroing "
                                                "+ backward, over the step's HIP-event time"),
        "counts": cnt,
        "bytes_model": "k_colors 224 N (16 B of mu + 192 B SH read, 16 B colour written); k_backward 4 R (n_comp) "
                       "+ 16 R_hit (dL/dRGBA) + 136 H_comp (id twice, 48-B record and 16-B colour in each of two "
                       "passes) + 112 P (one 56-B read-modify-write of the accumulators per distinct (16x16 block, "
                       "Gaussian) pair); k_finalize 64 N (accumulators) + 712 N_touched (record 48, SH 192, "
                       "236-B gradient row read and written); k_zero_f32 64 N; k_bw_cost 4 R + 4 B and k_bw_order 12 B "
                       "(the blocks' cost-first dispatch order; B = 16x16-pixel blocks)",
    }
    return {
        "workload": f"loss = sum (render - target)^2, target from means + N(0, 0.01); forward with checkpoint, "
                    f"dL/dRGBA, backward; {n} Gaussians, {W}x{H}, k {u.k}",
        "mrays": round(R / ms_fb / 1e3, 3), "ms_per_step": round(ms_fb, 3),
        "ratio_vs_forward": round(ms_fb / ms_fwd, 3),
        "ms_forward_with_checkpoint": round(ms_fwd_ck, 3), "ms_backward": round(ms_bwd, 3),
        "checkpoint_cap": cap, "checkpoint_bytes": R * cap * 4 + R * 4, "max_composited_per_ray": most,
        "roofline": roofline,
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
    ap.add_argument("--batch-alphas", type=float, nargs="+", default=[0.01, 0.1],
                    help="alpha_min values of the batched sweep")
    ap.add_argument("--batch-ks", type=int, nargs="+", default=[4, 8, 16, 32], help="k values of the batched sweep")
    ap.add_argument("--no-backward", action="store_true", help="skip the forward+backward measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the 3M-Gaussian 16-view forward+backward")
    ap.add_argument("--no-c4", action="store_true", help="skip the 200k-Gaussian 800x800 forward+backward")
    ap.add_argument("--no-train", action="store_true", help="skip the device build and training-step measurements")
    ap.add_argument("--cpu-sample", type=int, default=3, help="time 1/N of the rays on the CPU")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the launches whose HBM traffic roofline.traffic reports (for rocprofv3 --pmc)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 traffic measurement")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
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

    # HBM traffic first, by rocprofv3 child processes, before this process
    # initialises the GPU
    traffic, traffic_why = (None, "skipped (--no-traffic)")
    if not args.no_traffic:
        t_tr = time.time()
        traffic, traffic_why = measure_traffic(args)
        t_tr = time.time() - t_tr

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

    B_fwd = forward_bytes(R, st["h_eval"], st["h_comp"], raw.shape[0])
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
            "traffic": (traffic["FETCH_SIZE"]["k_forward_bytes"] + traffic["WRITE_SIZE"]["k_forward_bytes"]
                        if traffic else None),
            "traffic_detail": ({"fetch_bytes": traffic["FETCH_SIZE"]["k_forward_bytes"],
                                "write_bytes": traffic["WRITE_SIZE"]["k_forward_bytes"],
                                "fetch_scale": traffic["FETCH_SIZE"]["scale"],
                                "write_scale": traffic["WRITE_SIZE"]["scale"],
                                "traffic_over_algorithmic": round((traffic["FETCH_SIZE"]["k_forward_bytes"] +
                                                                   traffic["WRITE_SIZE"]["k_forward_bytes"]) / B_fwd, 4),
                                "kernel_trace_ms": traffic["kernel_ms"].get("k_forward"),
                                "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE child runs of this script "
                                          "(--traffic-probe: mean of the first two plain launches), scaled by the "
                                          "counters' reading for a 1 GiB device copy (--copy-calib)",
                                "measure_s": round(t_tr, 1)} if traffic else {"reason": traffic_why}),
            "algorithmic_bytes": int(B_fwd), "kernel": FWD_KERNEL,
            "measured_copy_gbs": round(copy_gbs, 1), "frac_of_measured_copy": round(achieved / copy_gbs, 5),
            # the bound the kernel meets first: VALU issue (no dense contraction,
            # so no MFMA; HBM traffic is a small fraction of its peak)
            "valu_issue": (dict(traffic["valu"], method="rocprofv3 --pmc " + " ".join(VALU_COUNTERS) +
                                " child run of this script (--traffic-probe: mean of the first two plain "
                                "launches); busy_frac = 4 SQ_ACTIVE_INST_VALU / (SIMDs x GRBM_GUI_ACTIVE "
                                "per XCD)")
                           if traffic and traffic.get("valu") else None),
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
        if traffic:
            # the backward's kernels, each from its own trace duration and measured traffic
            bb = out["fwd_bwd"]["roofline"]["backward"]["per_kernel_bytes"]
            per = {}
            for kname in ("k_colors", "k_backward", "k_finalize"):
                d = traffic["kernel_ms"].get(kname) or []
                ms_k = sum(d) / len(d) if d else None
                gbs = bb[kname] / (ms_k * 1e-3) / 1e9 if ms_k else None
                per[kname] = {"ms_trace": round(ms_k, 4) if ms_k else None, "algorithmic_bytes": bb[kname],
                              "achieved_gbs": round(gbs, 1) if gbs else None,
                              "frac": round(gbs / HBM_PEAK_GBS, 5) if gbs else None,
                              "fetch_bytes": traffic["FETCH_SIZE"]["backward_bytes"][kname],
                              "write_bytes": traffic["WRITE_SIZE"]["backward_bytes"][kname]}
            out["fwd_bwd"]["roofline"]["backward"]["per_kernel"] = per
    if not args.no_train:
        out["device_build"] = bench_device_build(gt, torch, dev, sp, raw, u, W, H, max(2, args.steps // 2))
        out["device_build"]["host_build_ms"] = round(t_build * 1e3, 1)
    if not args.no_batched:
        out["batched"] = bench_batched(gt, raw, scene, torch, dev, sp, args.batch_cams, tuple(args.batch_alphas),
                                       tuple(args.batch_ks))
    if not args.no_backward and not args.no_c4:
        out["fwd_bwd_c4"] = bench_c4(gt, torch, dev, sp, 5)
    if not args.no_backward and not args.no_train:
        out["train_c4"] = bench_train_c4(gt, torch, dev, sp, 10)
    if not args.no_c5:
        out["c5"] = bench_c5(gt, torch, dev, sp, 2)

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads, nproc, model, quota, aff = cpu_info()
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
            "nproc": nproc, "affinity_cpus": aff, "cgroup_cpu_quota": quota, "cpu_model": model,
            "sample": f"{n_s} seeded random rays (1/{args.cpu_sample} of the frame) of the same workload, "
                      f"oracle BVH variant (CPU restatement of the reference path, not the reference binary), "
                      f"{cs:.2f} s",
            "h_eval_per_ray": rr["stats"]["h_eval"] / n_s,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
