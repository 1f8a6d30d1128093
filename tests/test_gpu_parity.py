"""GPU parity of the fused HIP forward kernel against the CPU oracle.

Bar: per-ray, per-round hit lists (Gaussian ids), accept masks, round counts
and final T bit-exact; pixels within 1e-4 absolute per channel of the FP64
re-evaluation of the same hit lists; alpha channel exactly 1.
"""
import os

import numpy as np
import pytest

from helpers import check_parity, scene_pair, uniforms
import gtrace as gt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _print(*a):
    print(*a, flush=True)


def test_c1_config_full():
    """20k Gaussians uniform in a 4^3 box, 256x256, hFov 60, camera at (0,0,8),
    alpha_min 0.1, k 6, T_min 0.03 (the reference's CPU-runnable case)."""
    _print("case c1 full 256x256 k6")
    raw = gt.generate_raw(seed=0, config="c1")
    s, o = scene_pair(raw, 0.1)
    u = uniforms([0, 0, 8], [0, 0, 0], 256, 256, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    g, r = check_parity(s, o, u, 256, 256, label="c1")
    assert g["stats"]["h_comp"] > 1000


def test_cluster_scene_camera_inside():
    """Clustered scene (C2 generator at 1/10 size), camera inside the volume, k=16."""
    _print("case clusters 100k 160x90 k16")
    raw = gt.generate_raw(n=100000, seed=0, config="c2", n_clusters=8)
    s, o = scene_pair(raw, 0.01)
    c0 = raw[0::8, :3].mean(axis=0)  # centre of cluster 0 (Gaussian i is in cluster i mod 8)
    u = uniforms(list(c0 + np.array([0.4, -0.5, 2.5])), list(c0), 160, 90, hfov_deg=50, alpha_min=0.01, k=16,
                 T_min=0.001)
    g, r = check_parity(s, o, u, 160, 90, max_rounds=64, label="clusters")
    assert g["stats"]["rounds"] > g["stats"]["rays"]  # multi-round rays exist


@pytest.mark.parametrize("k,gpu_build", [(1, False), (3, False), (4, True), (16, False), (16, True)])
def test_exact_t_ties_ordered_by_id(k, gpu_build):
    """Every Gaussian three times, under shuffled ids: the copies have
    identical proxies, so each hit comes with two exact-t ties.  The
    k-buffer orders them by Gaussian id (the reference by its traversal
    order, which only a run of the reference could pin); ties fall inside
    rounds and across the k-th slot, where the relaunch at t_k + 1e-4 skips
    the remaining copies.  Hit lists bit-exact against the oracle (which
    orders by the same key) on the host-built and the GPU-built tree (the
    radix tree sees equal Morton codes)."""
    _print(f"case exact-t ties k={k} gpu_build={gpu_build}")
    base = gt.generate_raw(n=1500, seed=9, config="c1")
    n = base.shape[0]
    perm = np.random.default_rng(9).permutation(3 * n)
    raw = np.concatenate([base, base, base])[perm]
    orig = perm % n  # Gaussian id -> index of the copied Gaussian
    s = gt.Scene(raw, alpha_min=0.1, gpu_build=gpu_build)
    _, o = scene_pair(raw, 0.1, gpu=False)
    u = uniforms([0.2, -0.1, 8], [0, 0, 0], 64, 48, hfov_deg=50, alpha_min=0.1, k=k, T_min=0.01)
    g, _ = check_parity(s, o, u, 64, 48, max_rounds=256, label=f"ties k{k}")
    # the ties occur, and in id order
    h = g["hit_ids"].reshape(-1, k)
    a, b = h[:, :-1], h[:, 1:]
    valid = (a != gt.SENTINEL) & (b != gt.SENTINEL)
    tie = valid & (orig[np.where(valid, a, 0)] == orig[np.where(valid, b, 0)])
    if k > 1:
        assert tie.sum() > 100, f"only {tie.sum()} adjacent tied entries"
        assert np.all(a[tie] < b[tie])
    s.close()


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 13, 16, 24, 32])
def test_k_sweep(k):
    _print(f"case k sweep k={k}")
    raw = gt.generate_raw(n=20000, seed=3, config="c4")
    s, o = scene_pair(raw, 0.05)
    u = uniforms([0, 0, 3.0], [0, 0, 0], 48, 40, hfov_deg=50, alpha_min=0.05, k=k, T_min=0.01)
    check_parity(s, o, u, 48, 40, max_rounds=128, label=f"k{k}")


@pytest.mark.parametrize("W,H", [(37, 23), (1, 1), (8, 1), (1, 9), (65, 7)])
def test_odd_sizes(W, H):
    _print(f"case size {W}x{H}")
    raw = gt.generate_raw(n=5000, seed=4, config="c1")
    s, o = scene_pair(raw, 0.1)
    u = uniforms([0.3, 0.2, 5.0], [0, 0, 0], W, H, hfov_deg=40, alpha_min=0.1, k=6, T_min=0.03)
    check_parity(s, o, u, W, H, label=f"{W}x{H}")


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_degrees(deg):
    _print(f"case sh degree {deg}")
    raw = gt.generate_raw(n=5000, seed=5, config="c1")
    s, o = scene_pair(raw, 0.1)
    u = uniforms([0, 1, 6], [0, 0, 0], 32, 32, alpha_min=0.1, k=6, sh_degree_max=deg)
    check_parity(s, o, u, 32, 32, label=f"deg{deg}")


@pytest.mark.parametrize("early", [1, 2])
def test_early_stop(early):
    _print(f"case early stop {early}")
    raw = gt.generate_raw(n=20000, seed=6, config="c4")
    s, o = scene_pair(raw, 0.05)
    u = uniforms([0, 0, 3.0], [0, 0, 0], 32, 32, alpha_min=0.05, k=2, T_min=0.01, early_stop=early)
    g, _ = check_parity(s, o, u, 32, 32, label=f"early{early}")
    assert g["n_rounds"].max() <= early


@pytest.mark.parametrize("mode", [1, 2])
def test_debug_modes(mode):
    _print(f"case debug mode {mode}")
    raw = gt.generate_raw(n=20000, seed=7, config="c1")
    s, o = scene_pair(raw, 0.1)
    u = uniforms([0, 0, 8], [0, 0, 0], 64, 64, alpha_min=0.1, k=6, debug_mode=mode)
    W = H = 64
    g = s.render(u, W, H, max_rounds=1)
    import oracle as orc
    p = orc.params_from_uniforms(u, W, H)
    r = orc.render(o, p, max_rounds=1)
    np.testing.assert_array_equal(g["hit_ids"], r["hit_ids"])
    np.testing.assert_allclose(g["rgba"].reshape(-1, 4), r["rgba"], atol=1e-5, rtol=0)


def test_no_proxies_is_black():
    """No Gaussian above alpha_min: black image, A = 1, one empty round
    (the reference's build_tlas would panic here, lib.rs:518-519)."""
    _print("case empty scene")
    raw = gt.generate_raw(n=1000, seed=8, config="c1", logit_mean=-12.0, logit_std=0.1)
    s, o = scene_pair(raw, 0.1)
    assert s.info()["n_proxies"] == 0
    u = uniforms([0, 0, 8], [0, 0, 0], 16, 16, alpha_min=0.1, k=6)
    g, _ = check_parity(s, o, u, 16, 16, label="empty")
    assert np.all(g["rgba"][..., :3] == 0) and np.all(g["rgba"][..., 3] == 1)


def test_edge_hit_same_gaussian_twice():
    """A ray along +z through an axis-aligned Gaussian crosses the edge shared by
    two icosahedron faces: with inclusive barycentrics both faces report the hit,
    so the Gaussian occupies two k-buffer slots (reference behaviour)."""
    _print("case edge hit")
    raw = np.zeros((1, 59), np.float32)
    raw[0, 51] = 3.0          # opacity logit
    raw[0, 52:55] = np.log(0.5)
    raw[0, 55] = 1.0          # identity quaternion
    raw[0, 3:6] = 1.0
    s, o = scene_pair(raw, 0.01)
    u = uniforms([0, 0, 5.0], [0, 0, 0], 1, 1, alpha_min=0.01, k=4, T_min=1e-6)
    g, r = check_parity(s, o, u, 1, 1, label="edge")
    ids = g["hit_ids"][0, 0]
    assert list(ids[:2]) == [0, 0], ids


def test_pc_short_ply():
    """The reference's only data file: 100 real 3DGS Gaussians."""
    _print("case pc_short.ply")
    raw, deg = gt.read_ply(os.path.join(ROOT, "tests", "golden", "pc_short.ply"))
    assert deg == 3 and raw.shape[0] == 100
    c = raw[:, :3].mean(axis=0)
    s, o = scene_pair(raw, 0.01)
    u = uniforms(list(c + np.array([0.0, 0.0, 3.0])), list(c), 64, 64, alpha_min=0.01, k=4, T_min=0.001)
    check_parity(s, o, u, 64, 64, label="pc_short")


def test_determinism_same_frame_twice():
    _print("case determinism")
    raw = gt.generate_raw(n=50000, seed=9, config="c2")
    s = gt.Scene(raw, alpha_min=0.01)
    u = uniforms([0.5, 0.5, 2.0], [0, 0, 0], 96, 64, hfov_deg=50, alpha_min=0.01, k=16, T_min=0.001)
    a = s.render(u, 96, 64, want_T=True)
    b = s.render(u, 96, 64, want_T=True)
    assert a["rgba"].tobytes() == b["rgba"].tobytes() and a["T"].tobytes() == b["T"].tobytes()


@pytest.mark.parametrize("k", [1, 6, 16, 32])
def test_packet_traversal_parity(k):
    """Wave-wide packet traversal (tuning path 3) against the oracle: the
    clustered scene with the camera inside, and the C1 scene."""
    _print(f"case packet traversal k{k}")
    raw = gt.generate_raw(n=100000, seed=0, config="c2", n_clusters=8)
    s, o = scene_pair(raw, 0.01)
    s.set_tuning(path=3)
    c0 = raw[0::8, :3].mean(axis=0)
    u = uniforms(list(c0 + np.array([0.3, 0.2, 1.5])), list(c0), 96, 64, hfov_deg=50, alpha_min=0.01, k=k,
                 T_min=1e-3)
    check_parity(s, o, u, 96, 64, max_rounds=max(16, 512 // k), label=f"packet clusters k{k}")
    raw = gt.generate_raw(seed=0, config="c1")
    s, o = scene_pair(raw, 0.1)
    s.set_tuning(path=3, tile_w=16)
    u = uniforms([0, 0, 8], [0, 0, 0], 80, 72, hfov_deg=60, alpha_min=0.1, k=k, T_min=0.03)
    check_parity(s, o, u, 80, 72, max_rounds=max(16, 512 // k), label=f"packet c1 k{k}")


@pytest.mark.parametrize("k", [1, 16])
def test_persistent_packet_kernel_parity(k):
    """The persistent packet kernel (tuning path 4: resident waves take
    64-ray tiles from per-XCD queues) against the oracle, with a grid small
    enough that every wave renders many tiles and takes tiles from the other
    XCDs' ranges."""
    _print(f"case persistent packet kernel k{k}")
    raw = gt.generate_raw(n=100000, seed=0, config="c2", n_clusters=8)
    s, o = scene_pair(raw, 0.01)
    c0 = raw[0::8, :3].mean(axis=0)
    u = uniforms(list(c0 + np.array([0.3, 0.2, 1.5])), list(c0), 123, 77, hfov_deg=50, alpha_min=0.01, k=k,
                 T_min=1e-3)
    for bpc in (1, 4):
        s.set_tuning(path=4, blocks_per_cu=bpc)
        check_parity(s, o, u, 123, 77, max_rounds=max(16, 512 // k), label=f"persistent k{k} bpc{bpc}")
    for mode in (1, 2):
        u.debug_mode = mode
        g = s.render(u, 123, 77, max_rounds=1)
        import oracle as orc
        r = orc.render(o, orc.params_from_uniforms(u, 123, 77), max_rounds=1)
        np.testing.assert_array_equal(g["hit_ids"], r["hit_ids"])
        np.testing.assert_allclose(g["rgba"].reshape(-1, 4), r["rgba"], atol=1e-5, rtol=0)


@pytest.mark.parametrize("mode", [1, 2])
def test_packet_traversal_debug_modes(mode):
    _print(f"case packet debug mode {mode}")
    raw = gt.generate_raw(n=20000, seed=7, config="c1")
    s, o = scene_pair(raw, 0.1)
    s.set_tuning(path=3)
    u = uniforms([0, 0, 8], [0, 0, 0], 64, 64, alpha_min=0.1, k=6, debug_mode=mode)
    g = s.render(u, 64, 64, max_rounds=1)
    import oracle as orc
    r = orc.render(o, orc.params_from_uniforms(u, 64, 64), max_rounds=1)
    np.testing.assert_array_equal(g["hit_ids"], r["hit_ids"])
    np.testing.assert_allclose(g["rgba"].reshape(-1, 4), r["rgba"], atol=1e-5, rtol=0)


@pytest.mark.parametrize("tile_w", [0, 4, 16, 32])
def test_host_image_equals_device_image(tile_w):
    """A host-image render equals, byte for byte, a device-buffer render and a
    host render that also returns counters, for every ray tile shape and
    debug mode, at sizes that leave partial tiles."""
    import torch
    _print(f"case host image vs device image tile_w={tile_w}")
    raw = gt.generate_raw(n=5000, seed=4, config="c1")
    s = gt.Scene(raw, alpha_min=0.1)
    s.set_tuning(tile_w=tile_w)
    dev = torch.device("cuda", 0)
    for (W, H) in [(1, 1), (65, 7), (37, 23), (300, 170), (1000, 563)]:
        for mode in (0, 1, 2):
            u = uniforms([0.3, 0.2, 5.0], [0, 0, 0], W, H, hfov_deg=40, alpha_min=0.1, k=6, T_min=0.03)
            u.debug_mode = mode
            banded = s.render(u, W, H)["rgba"]
            single = s.render(u, W, H, want_stats=True)["rgba"]
            img = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
            s.render_device(u, W, H, img.data_ptr())
            torch.cuda.synchronize()
            d = img.cpu().numpy()
            assert np.array_equal(banded.view(np.uint32), single.view(np.uint32)), f"{W}x{H} mode {mode}"
            assert np.array_equal(banded.view(np.uint32), d.view(np.uint32)), f"{W}x{H} mode {mode}"
    s.close()
