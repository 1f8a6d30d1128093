"""GPU parity of the packed-record entry point gt_scene_create_packed.

It takes the reference's gsData.gaussiansBuffer unchanged: n records of
320 B as src/GS_ply.ts:198-253 packs them (position, exp'd scale,
normalised quaternion, sigmoid opacity, 16 SH vec3).  The buffer does not
carry the JS doubles that gaussianSceneToBvh reads (src/main.ts:73-78), so
proxies are built from its f32 values; the oracle restates exactly that
(gto_build_proxies_packed, pinned by the proxy tangency check).  Bar: the
scene's records equal the buffer bit for bit, its proxies equal the
oracle's, and renders meet the parity bar against the oracle on the same
records.
"""
import os

import numpy as np
import pytest

import gtrace as gt
import oracle as orc
from helpers import check_parity, uniforms

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _print(*a):
    print(*a, flush=True)


def _case(name):
    if name == "pc_short":
        raw, _ = gt.read_ply(os.path.join(ROOT, "tests", "golden", "pc_short.ply"))
        c = raw[:, :3].mean(axis=0)
        return raw, 0.01, (list(c + np.array([0.0, 0.0, 3.0])), list(c)), 64, 64, dict(k=4, T_min=1e-3, hfov_deg=60)
    raw = gt.generate_raw(seed=0, config="c1")
    return raw, 0.1, ([0.0, 0.0, 8.0], [0.0, 0.0, 0.0]), 256, 256, dict(k=6, T_min=0.03, hfov_deg=60)


@pytest.mark.parametrize("name", ["pc_short", "c1"])
def test_packed_entry_parity(name):
    _print(f"case packed records {name}")
    raw, alpha, cam, W, H, kw = _case(name)
    packed = orc.OracleScene(raw, alpha).packed  # GS_ply.ts's 320-B records of these PLY values
    s = gt.Scene(packed=packed, alpha_min=alpha)
    o = orc.OracleScene(packed=packed, alpha_min=alpha)
    hd = s.host_data()
    assert np.array_equal(hd["packed"].view(np.uint32), packed.view(np.uint32))
    assert np.array_equal(hd["proxy_ids"], o.proxy_ids)
    assert np.array_equal(hd["o2w"].view(np.uint32), o.o2w.view(np.uint32))
    assert np.array_equal(hd["w2o"].view(np.uint32), o.w2o.view(np.uint32))
    u = uniforms(cam[0], cam[1], W, H, alpha_min=alpha, **kw)
    g, r = check_parity(s, o, u, W, H, max_rounds=48, label=f"packed {name}")
    assert g["stats"]["h_comp"] > (20 if name == "pc_short" else 1000)
    # the buffer as raw bytes (an ArrayBuffer handed over unchanged) renders the same
    s2 = gt.Scene(packed=np.frombuffer(packed.tobytes(), dtype=np.float32), alpha_min=alpha)
    g2 = s2.render(u, W, H, max_rounds=48)
    assert g2["rgba"].tobytes() == g["rgba"].tobytes() and np.array_equal(g2["hit_ids"], g["hit_ids"])
    # a new alpha_min re-filters and rebuilds from the same f32 values
    a2 = alpha * 2.5
    s.set_alpha_min(a2)
    o2 = orc.OracleScene(packed=packed, alpha_min=a2)
    u2 = uniforms(cam[0], cam[1], W, H, alpha_min=a2, **kw)
    check_parity(s, o2, u2, W, H, max_rounds=48, label=f"packed {name} alpha {a2}")
    s.close()
    s2.close()
