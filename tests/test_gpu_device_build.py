"""GPU scene build (gt_scene_create_device / gt_scene_update_device).

Bar: the device-built records, proxy ids and proxy transforms equal the host
builder's bit for bit; renders of a device-built scene (GPU radix-tree BVH)
equal those of the host-built scene (binned-SAH BVH) and the oracle: hit
lists, accept masks, round counts and T bit-exact, pixels within 1e-4 of
FP64.  The BVH never changes which hits a ray finds.
"""
import numpy as np
import pytest
import torch

from helpers import check_parity, uniforms
import gtrace as gt
import oracle as orc

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _dev_scene(raw, alpha_min):
    t = torch.from_numpy(np.ascontiguousarray(raw, np.float32)).to("cuda:0")
    s = gt.Scene(raw_dev=t.data_ptr(), n=raw.shape[0], alpha_min=alpha_min)
    return s, t


def _same_render(a, b, u, W, H, max_rounds=64, label=""):
    ga = a.render(u, W, H, max_rounds=max_rounds, want_stats=True)
    gb = b.render(u, W, H, max_rounds=max_rounds, want_stats=True)
    for key in ("hit_ids", "accept", "n_rounds"):
        np.testing.assert_array_equal(ga[key], gb[key], err_msg=f"{label}: {key}")
    np.testing.assert_array_equal(ga["T"].view(np.uint32), gb["T"].view(np.uint32), err_msg=f"{label}: T")
    np.testing.assert_array_equal(ga["rgba"].view(np.uint32), gb["rgba"].view(np.uint32), err_msg=f"{label}: rgba")
    # "hits", "proxies" and "nodes" count traversal work, which depends on the tree
    for key in ("rays", "rounds", "h_eval", "h_comp"):
        assert ga["stats"][key] == gb["stats"][key], (label, key)
    return ga, gb


@pytest.mark.parametrize("cfg,n,alpha", [("c1", 3000, 0.1), ("c2", 40000, 0.01), ("c4", 5000, 0.2)])
def test_device_records_equal_host(cfg, n, alpha):
    _print(f"case device records {cfg} n={n} alpha={alpha}")
    raw = gt.generate_raw(n=n, seed=21, config=cfg)
    h = gt.Scene(raw, alpha_min=alpha)
    d, _t = _dev_scene(raw, alpha)
    hd, dd = h.host_data(), d.host_data()
    for key in ("packed", "o2w", "w2o"):
        np.testing.assert_array_equal(hd[key].view(np.uint32), dd[key].view(np.uint32), err_msg=key)
    np.testing.assert_array_equal(hd["proxy_ids"], dd["proxy_ids"])
    # and the oracle's restatement of the same stage
    packed, act = orc.activate(raw)
    ids, o2w, w2o = orc.build_proxies(raw, act, alpha)
    assert np.array_equal(dd["packed"].view(np.uint32), packed.view(np.uint32))
    assert np.array_equal(dd["proxy_ids"], ids)
    assert np.array_equal(dd["w2o"].view(np.uint32), w2o.view(np.uint32))
    info = d.info()
    assert info["n_proxies"] == len(ids) and 0 < info["max_stack"] <= 96 and info["bvh_depth"] >= 1


def test_device_scene_c1_parity_and_equals_host():
    """C1 (20k Gaussians, 256x256, k 6): device-built == host-built == oracle."""
    _print("case device c1 256x256 k6")
    raw = gt.generate_raw(seed=0, config="c1")
    h = gt.Scene(raw, alpha_min=0.1)
    d, _t = _dev_scene(raw, 0.1)
    o = orc.OracleScene(raw, 0.1)
    u = uniforms([0, 0, 8], [0, 0, 0], 256, 256, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    check_parity(d, o, u, 256, 256, label="device c1")
    _same_render(h, d, u, 256, 256, label="c1 host vs device")


@pytest.mark.parametrize("k", [1, 7, 16, 32])
def test_device_scene_clusters(k):
    """Clustered scene, camera inside the volume, several k; both traversal kernels."""
    _print(f"case device clusters k={k}")
    raw = gt.generate_raw(n=60000, seed=5, config="c2", n_clusters=8)
    h = gt.Scene(raw, alpha_min=0.01)
    d, _t = _dev_scene(raw, 0.01)
    c0 = raw[0::8, :3].mean(axis=0)
    u = uniforms(list(c0 + np.array([0.4, -0.5, 2.5])), list(c0), 96, 64, hfov_deg=50, alpha_min=0.01, k=k,
                 T_min=0.001)
    ref, _ = _same_render(h, d, u, 96, 64, max_rounds=160, label=f"k{k}")
    d.set_tuning(path=1)  # per-lane traversal on the device-built tree
    g = d.render(u, 96, 64, max_rounds=160)
    np.testing.assert_array_equal(g["hit_ids"], ref["hit_ids"])
    np.testing.assert_array_equal(g["rgba"].view(np.uint32), ref["rgba"].view(np.uint32))


@pytest.mark.parametrize("n,alpha", [(1, 0.01), (9, 0.01), (300, 0.999), (0, 0.1)])
def test_device_small_and_empty(n, alpha):
    """One proxy, just above one leaf, no proxy at all (alpha_min above every opacity), no Gaussians."""
    _print(f"case device small n={n} alpha={alpha}")
    raw = gt.generate_raw(n=max(n, 1), seed=8, config="c4", logit_mean=2.0, log_scale_mean=-1.0)[:n]
    o = orc.OracleScene(raw, alpha) if n else None
    if n == 0:
        t = torch.zeros(1, gt.GT_RAW_FLOATS, device="cuda:0")
        d = gt.Scene(raw_dev=t.data_ptr(), n=0, alpha_min=alpha)
    else:
        d, t = _dev_scene(raw, alpha)
    u = uniforms([0, 0, 4.0], [0, 0, 0], 40, 30, hfov_deg=60, alpha_min=alpha, k=4, T_min=0.01)
    if n:
        check_parity(d, o, u, 40, 30, label=f"n{n}")
    g = d.render(u, 40, 30, max_rounds=4)
    if d.info()["n_proxies"] == 0:
        assert np.all(g["rgba"][..., :3] == 0.0) and np.all(g["hit_ids"] == gt.SENTINEL)


def test_device_coincident_means():
    """Many Gaussians at exactly the same position: equal Morton codes, split by id."""
    _print("case device coincident means")
    raw = gt.generate_raw(n=4000, seed=9, config="c4")
    raw[:3000, :3] = raw[0, :3]
    d, _t = _dev_scene(raw, 0.05)
    o = orc.OracleScene(raw, 0.05)
    c = [float(x) for x in raw[0, :3]]
    u = uniforms([c[0], c[1], c[2] + 2.0], c, 48, 48, hfov_deg=40, alpha_min=0.05, k=8, T_min=0.001)
    check_parity(d, o, u, 48, 48, max_rounds=512, label="coincident")
    assert d.info()["max_stack"] <= 96


def test_update_device_and_alpha_min():
    """After a parameter update the rebuilt scene equals a fresh host-built scene,
    and set_alpha_min rebuilds on the GPU."""
    _print("case update_device")
    raw = gt.generate_raw(n=30000, seed=11, config="c4")
    d, t = _dev_scene(raw, 0.02)
    u = uniforms([0, 0.3, 3.0], [0, 0, 0], 64, 48, hfov_deg=50, alpha_min=0.02, k=12, T_min=0.001)
    raw2 = raw.copy()
    rng = np.random.default_rng(3)
    raw2[:, :3] += rng.normal(0, 0.05, (raw.shape[0], 3)).astype(np.float32)
    raw2[:, 51] += rng.normal(0, 0.5, raw.shape[0]).astype(np.float32)
    t2 = torch.from_numpy(raw2).to("cuda:0")
    d.update_device(t2.data_ptr())
    _same_render(gt.Scene(raw2, alpha_min=0.02), d, u, 64, 48, max_rounds=128, label="updated")
    d.set_alpha_min(0.2)
    h = gt.Scene(raw2, alpha_min=0.2)
    assert d.info()["n_proxies"] == h.info()["n_proxies"]
    _same_render(h, d, u, 64, 48, max_rounds=128, label="alpha 0.2")


def test_device_scene_backward_equals_host():
    """Gradients of a device-built scene equal the host-built scene's (FP32
    atomics: summation order differs, so within 1e-5 relative)."""
    _print("case device backward")
    raw = gt.generate_raw(n=20000, seed=12, config="c4")
    h = gt.Scene(raw, alpha_min=0.01)
    d, _t = _dev_scene(raw, 0.01)
    W, H = 64, 64
    u = uniforms([0, 0, 3.0], [0, 0, 0], W, H, hfov_deg=50, alpha_min=0.01, k=16, T_min=0.001)
    dl = torch.randn(H, W, 4, device="cuda:0")
    out = []
    for s in (h, d):
        g = torch.zeros(raw.shape[0], gt.GT_RAW_FLOATS, device="cuda:0")
        gt.backward(s, u, W, H, dl.data_ptr(), g.data_ptr())
        torch.cuda.synchronize()
        out.append(g.cpu().numpy().astype(np.float64))
    scale = np.abs(out[0]).max()
    assert scale > 0
    assert np.abs(out[0] - out[1]).max() <= 1e-5 * scale


def test_host_input_gpu_build():
    """gt_scene_create_raw_gpu: host parameters uploaded and built on the GPU."""
    _print("case create_raw_gpu")
    raw = gt.generate_raw(n=20000, seed=14, config="c1")
    h = gt.Scene(raw, alpha_min=0.1)
    g = gt.Scene(raw, alpha_min=0.1, gpu_build=True)
    u = uniforms([0.5, 0.2, 7.0], [0, 0, 0], 96, 80, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03)
    _same_render(h, g, u, 96, 80, label="raw_gpu")
    x = torch.from_numpy(raw).to("cuda:0")
    g.update_device(x.data_ptr())  # a scene from host input can be updated from device parameters
    _same_render(h, g, u, 96, 80, label="raw_gpu updated")


def _bad_raws():
    base = gt.generate_raw(n=4000, seed=3, config="c1")
    base[1737, 51] = 3.0  # opacity 0.95: a proxy at alpha_min 0.05
    cases = []
    for label, col, val in [("nan mean", 1, np.nan), ("scale overflows f32", 52, 100.0),
                            ("position near FLT_MAX", 0, 3.0e38), ("scale x proxy factor overflows", 53, 88.5),
                            ("inverse overflows", 52, -87.0)]:
        r = base.copy()
        r[1737, col] = val
        cases.append((label, r))
    r = base.copy()
    r[1737, 55:59] = 0.0
    cases.append(("zero quaternion", r))
    return cases


@pytest.mark.parametrize("label,raw", _bad_raws(), ids=[c[0] for c in _bad_raws()])
def test_invalid_records_rejected_by_both_builders(label, raw):
    """The host and the GPU builder refuse the same records (non-finite
    parameters, zero quaternion, scales 0 or inf in f32, proxies whose
    transform, inverse or box overflows f32) with GT_E_INVALID naming the
    Gaussian; gt_scene_update_device and gt_scene_refit_device refuse them
    for a valid scene, which renders as before once updated with valid
    parameters again."""
    _print("invalid record:", label)
    with pytest.raises(gt.GTError) as e:
        gt.Scene(raw=raw, alpha_min=0.05)
    assert e.value.code == -1 and "Gaussian 1737" in str(e.value), str(e.value)
    t = torch.from_numpy(np.ascontiguousarray(raw, np.float32)).to("cuda:0")
    with pytest.raises(gt.GTError) as e:
        gt.Scene(raw_dev=t.data_ptr(), n=raw.shape[0], alpha_min=0.05)
    assert e.value.code == -1 and "Gaussian 1737" in str(e.value), str(e.value)
    good = gt.generate_raw(n=4000, seed=3, config="c1")
    s, tg = _dev_scene(good, 0.05)
    u = uniforms([0, 0.3, 4.0], [0, 0, 0], 64, 48, hfov_deg=50, alpha_min=0.05, k=8, T_min=0.001)
    before = s.render(u, 64, 48)["rgba"]
    with pytest.raises(gt.GTError) as e:
        s.update_device(t.data_ptr())
    assert e.value.code == -1 and "Gaussian 1737" in str(e.value), str(e.value)
    s.update_device(tg.data_ptr())
    np.testing.assert_array_equal(s.render(u, 64, 48)["rgba"], before)
    h = gt.Scene(raw=good, alpha_min=0.05)
    with pytest.raises(gt.GTError) as e:
        h.refit_device(t.data_ptr())
    assert e.value.code == -1 and "Gaussian 1737" in str(e.value), str(e.value)
    h.refit_device(tg.data_ptr())
    np.testing.assert_array_equal(h.render(u, 64, 48)["rgba"], before)
