"""Shared scene/camera builders and parity assertions for the tests."""
import numpy as np

import gtrace as gt
import oracle as orc

PIXEL_TOL = 1e-4  # absolute per channel, FP32 kernel vs FP64 re-evaluation of the same hit lists


def scene_pair(raw, alpha_min, gpu=True):
    """(library scene or None, oracle scene) over the same PLY-raw parameters."""
    s = gt.Scene(raw, alpha_min=alpha_min) if gpu else None
    o = orc.OracleScene(raw, alpha_min)
    return s, o


def uniforms(pos, target, W, H, **kw):
    tw = gt.camera_look_at(pos, target)
    return gt.make_uniforms(tw, W, H, **kw)


def render_rects_both(s, u, W, H, max_rounds):
    """The frame rendered with screen-rectangle culling and precomputed
    per-Gaussian colours forced on (tuning rects=1, colors=1) and both off
    (rects=2, colors=2: every composited hit evaluates its SH record), the
    scene's tuning otherwise kept: outputs identical bit for bit, the same
    hits and node visits, no more proxy tests with the culling.  Returns the
    render with both on."""
    t0 = s.get_tuning()
    try:
        s.set_tuning(**dict(t0, rects=1, colors=1))
        g = s.render(u, W, H, max_rounds=max_rounds, want_stats=True)
        s.set_tuning(**dict(t0, rects=2, colors=2))
        g2 = s.render(u, W, H, max_rounds=max_rounds, want_stats=True)
    finally:
        s.set_tuning(**t0)
    np.testing.assert_array_equal(g["rgba"].view(np.uint32), g2["rgba"].view(np.uint32), err_msg="rects/colors: rgba")
    for key in ("hit_ids", "accept", "n_rounds"):
        np.testing.assert_array_equal(g[key], g2[key], err_msg=f"rects/colors: {key}")
    np.testing.assert_array_equal(g["T"].view(np.uint32), g2["T"].view(np.uint32), err_msg="rects/colors: T")
    for key in ("rays", "rounds", "h_eval", "h_comp", "hits", "nodes"):
        assert g["stats"][key] == g2["stats"][key], f"rects/colors: stats {key}"
    assert g["stats"]["proxies"] <= g2["stats"]["proxies"]
    return g


def check_parity(s, o, u, W, H, max_rounds=24, label=""):
    """Kernel vs oracle: hit lists, accept masks, rounds and T bit-exact;
    pixels within PIXEL_TOL of the FP64 re-evaluation.  The kernel renders
    with and without screen-rectangle culling and precomputed colours
    (identical outputs required); the render with both is checked.  Returns
    both results."""
    g = render_rects_both(s, u, W, H, max_rounds)
    p = orc.params_from_uniforms(u, W, H)
    r = orc.render(o, p, max_rounds=max_rounds)
    R = W * H
    assert r["stats"]["overflow"] == 0, f"{label}: raise max_rounds"
    np.testing.assert_array_equal(g["n_rounds"], r["n_rounds"], err_msg=f"{label}: rounds")
    bad = np.nonzero((g["hit_ids"] != r["hit_ids"]).any(axis=(1, 2)))[0]
    assert bad.size == 0, f"{label}: hit lists differ on {bad.size} rays, first ray {bad[:5]}"
    np.testing.assert_array_equal(g["accept"], r["accept"], err_msg=f"{label}: accept masks")
    np.testing.assert_array_equal(g["T"].view(np.uint32), r["T"].view(np.uint32), err_msg=f"{label}: T bits")
    rgb64, _ = orc.reeval_f64(o, p, g["n_rounds"], g["hit_ids"], g["accept"])
    rgba = g["rgba"].reshape(R, 4)
    err = np.abs(rgba[:, :3].astype(np.float64) - rgb64).max() if R else 0.0
    assert err <= PIXEL_TOL, f"{label}: pixel error {err} > {PIXEL_TOL}"
    assert np.all(rgba[:, 3] == 1.0)
    st = g["stats"]
    assert st["rays"] == R and st["rounds"] == int(r["n_rounds"].sum())
    assert st["h_eval"] == r["stats"]["h_eval"] and st["h_comp"] == r["stats"]["h_comp"]
    return g, r
