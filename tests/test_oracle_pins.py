"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors, known-answer tests or fixtures beyond
data/pc_short.ply (its only unit test, webrtx/bvh/src/lib.rs:605-749, has no
assertion).  Each part of the oracle is therefore checked against an
independent derivation: closed forms, FP64 re-derivations with textbook
formulas (Moller-Trumbore, real spherical harmonics from scipy, the
line-quadric minimum), brute force on tiny inputs and invariants.
test_mutation_detected plants plausible mistakes in the oracle (through its
mutation switch) and requires the matching pin to fail.
"""
import math
import os

import numpy as np
import pytest
import scipy.special as sps

import gtrace as gt
import oracle as orc

F32 = np.float32


# ---------------------------------------------------------------- helpers
def quat_to_R(q):
    """Standard rotation matrix of a unit quaternion (w, x, y, z)."""
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def ico_geometry():
    """Icosahedron of src/main.ts:12-25 from its defining formulas, in double
    over the f32-rounded coordinates the reference stores."""
    r = (3 + math.sqrt(5.0)) / (2 * math.sqrt(3.0))
    s = float(F32(1 / r))
    t = float(F32((1.0 + math.sqrt(5.0)) / (2.0 * r)))
    V = np.array([[-s, t, 0], [s, t, 0], [-s, -t, 0], [s, -t, 0], [0, -s, t], [0, s, t], [0, -s, -t], [0, s, -t],
                  [t, 0, -s], [t, 0, s], [-t, 0, -s], [-t, 0, s]], np.float64)
    I = np.array([0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                  3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]).reshape(20, 3)
    return V, I


def packed_record(mu, log_scale, quat, logit, dc=(0, 0, 0), rest=None):
    raw = np.zeros((1, 59), F32)
    raw[0, :3] = mu
    raw[0, 3:6] = dc
    if rest is not None:
        raw[0, 6:51] = rest
    raw[0, 51] = logit
    raw[0, 52:55] = log_scale
    raw[0, 55:59] = quat
    return raw


def mt64(o, d, p0, p1, p2):
    """Moller-Trumbore in double: (t, u, v) of the line o + t d with the triangle."""
    e1, e2 = p1 - p0, p2 - p0
    h = np.cross(d, e2)
    a = e1 @ h
    f = 1.0 / a
    s = o - p0
    u = f * (s @ h)
    q = np.cross(s, e1)
    v = f * (d @ q)
    t = f * (e2 @ q)
    return t, u, v


# ---------------------------------------------------------------- pins
def check_exp64():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.linspace(-699.0, 699.0, 60001), -np.logspace(-12, 2.5, 3000),
                         np.logspace(-12, 2.5, 3000), rng.uniform(-40, 5, 20000), [0.0, -1e-300, 1e-300]])
    got = np.array([orc.exp64(float(x)) for x in xs])
    ref = np.exp(xs)
    rel = np.abs(got - ref) / ref
    assert rel.max() < 2e-15, f"exp64 relative error {rel.max()} at x={xs[rel.argmax()]}"
    assert orc.exp64(-700.5) == 0.0 and orc.exp64(-1e4) == 0.0 and orc.exp64(float("nan")) == 0.0
    assert math.isinf(orc.exp64(700.0))
    assert orc.exp64(1.0) == math.e or abs(orc.exp64(1.0) / math.e - 1) < 3e-16


def check_log64():
    """gt_log64 (the activations' ln: proxy scale sqrt(2 ln(opacity / alpha_min)),
    src/main.ts:84) against numpy's log over the double range."""
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.logspace(-300, 300, 20001), 1.0 + rng.uniform(-1e-3, 1e-3, 20000),
                         rng.uniform(1.0, 200.0, 20000), 1.0 + np.logspace(-15, -1, 2000),
                         [5e-324, 2.2250738585072014e-308, 0.5, 2.0, 1.4142135623730951, 0.7071067811865476]])
    got = np.array([orc.log64(float(x)) for x in xs])
    ref = np.log(xs)
    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    assert rel.max() < 6e-16, f"log64 relative error {rel.max()} at x={xs[nz][rel.argmax()]}"
    assert orc.log64(1.0) == 0.0
    assert abs(orc.log64(2.0) - math.log(2.0)) <= 1.2e-16
    assert math.isnan(orc.log64(0.0)) and math.isnan(orc.log64(-1.0)) and math.isnan(orc.log64(float("nan")))
    assert math.isinf(orc.log64(float("inf")))


def check_icosahedron():
    V, I = ico_geometry()
    for f in range(20):
        p0, p1, p2 = V[I[f]]
        n = np.cross(p0 - p2, p1 - p0)  # cross(e1, e0) of intersect.glsl:78-80
        nh = n / np.linalg.norm(n)
        assert abs(nh @ p0 - 1.0) < 1e-6, "face plane at distance 1 (inradius 1)"
        assert nh @ (p0 + p1 + p2) > 0, "faces are outward with n = cross(e1, e0)"
    assert np.allclose(np.linalg.norm(V, axis=1), 1.2584085723648188, atol=1e-6)


def check_ray_dir():
    pos, tgt = [3.0, 2.0, 7.0], [0.2, -0.1, 0.3]
    tw = gt.camera_look_at(pos, tgt)
    W, H = 33, 17
    u = gt.make_uniforms(tw, W, H, hfov_deg=50)
    p = orc.params_from_uniforms(u, W, H)
    M = tw.astype(np.float64).reshape(4, 4)  # row c = column c of the column-major matrix
    X, Y, Z = -M[0, :3], M[1, :3], -M[2, :3]
    ax = math.tan(0.5 * 50 * math.pi / 180)
    ay = float(F32(ax)) * H / W
    for px, py in [(0, 0), (32, 16), (5, 11), (16, 8), (31, 2)]:
        ndc = np.array([(px + 0.5) / W * 2 - 1, (py + 0.5) / H * 2 - 1])
        v = ndc[0] * ax * X + ndc[1] * ay * Y + Z
        v /= np.linalg.norm(v)
        d = orc.ray_dir(p, px, py)
        assert np.abs(d - v).max() < 2e-6, (px, py, d, v)
    # centre pixel looks along -back = -(pos - target)/|pos - target|
    b = np.array(pos) - np.array(tgt)
    assert np.abs(orc.ray_dir(p, 16, 8) + b / np.linalg.norm(b)).max() < 2e-6
    # right/up of the camera.ts matrix: x grows to -col0 ... the image is rotated 180 deg vs y-down
    d00, dWH = orc.ray_dir(p, 0, 0), orc.ray_dir(p, W - 1, H - 1)
    assert np.abs((d00 + dWH) / 2 - (d00 + dWH) @ Z / 2 * Z / (Z @ Z)).max() < 1e-6


def _reference_pixel_offset(px, py, W, frames):
    """GS_shader.ts:8-28, 281-283 restated in Python integers: seed =
    getNewSeed(W*py + px, frames, 8) (TEA rounds, first word), then two
    rng draws (LCG, low 24 bits / 2^24)."""
    m = 0xFFFFFFFF
    v0, v1, s0 = (W * py + px) & m, frames & m, 0
    for _ in range(8):
        s0 = (s0 + 0x9E3779B9) & m
        v0 = (v0 + ((((v1 << 4) & m) + 0xA341316C) & m ^ ((v1 + s0) & m) ^ (((v1 >> 5) + 0xC8013EA4) & m))) & m
        v1 = (v1 + ((((v0 << 4) & m) + 0xAD90777D) & m ^ ((v0 + s0) & m) ^ (((v0 >> 5) + 0x7E95761E) & m))) & m
    seed, offs = v0, []
    for _ in range(2):
        seed = (1664525 * seed + 1013904223) & m
        offs.append((seed & 0x00FFFFFF) / float(0x01000000))
    return offs


def check_aa_jitter():
    """Antialiasing pixel offsets (compiled out in the reference): the oracle's
    jittered ray equals the camera ray through (px + ox, py + oy) with the
    offsets of an independent restatement of getNewSeed/rng; offsets change
    with the frame index and are uniform on [0, 1)."""
    pos, tgt = [3.0, 2.0, 7.0], [0.2, -0.1, 0.3]
    tw = gt.camera_look_at(pos, tgt)
    W, H = 33, 17
    M = tw.astype(np.float64).reshape(4, 4)
    X, Y, Z = -M[0, :3], M[1, :3], -M[2, :3]
    ax = math.tan(0.5 * 50 * math.pi / 180)
    ay = float(F32(ax)) * H / W
    for frames in (0, 1, 7, 123456):
        u = gt.make_uniforms(tw, W, H, hfov_deg=50, accumulated_frames=frames)
        p = orc.params_from_uniforms(u, W, H, antialias=True)
        for px, py in [(0, 0), (32, 16), (5, 11), (16, 8)]:
            ox, oy = _reference_pixel_offset(px, py, W, frames)
            ndc = np.array([(px + ox) / W * 2 - 1, (py + oy) / H * 2 - 1])
            v = ndc[0] * ax * X + ndc[1] * ay * Y + Z
            v /= np.linalg.norm(v)
            d = orc.ray_dir(p, px, py)
            assert np.abs(d - v).max() < 2e-6, (frames, px, py, d, v)
    offs = np.array([_reference_pixel_offset(i % 97, i // 97, 97, 3) for i in range(97 * 40)])
    assert offs.min() >= 0.0 and offs.max() < 1.0 and abs(offs.mean() - 0.5) < 0.02
    assert _reference_pixel_offset(5, 3, 33, 1) != _reference_pixel_offset(5, 3, 33, 2)


def check_proxy_tangent(packed_input=False):
    """Every proxy face is tangent to the Mahalanobis sphere of radius
    s = sqrt(2 ln(opacity / alpha_min)) around the mean (main.ts:84-87),
    with Sigma = R S^2 R^T, R the standard rotation of the quaternion.
    packed_input: the proxies of a packed 320-B buffer, built from its f32
    opacity, scale and quaternion (the JS doubles are not in the buffer)."""
    raw = gt.generate_raw(n=200, seed=11, config="c1")
    alpha = 0.05
    packed, act = orc.activate(raw)
    if packed_input:
        act, ids, o2w, w2o = orc.build_proxies_packed(packed, alpha)
        assert np.array_equal(act[:, 1:4], packed[:, 4:7].astype(np.float64))
    else:
        ids, o2w, w2o = orc.build_proxies(raw, act, alpha)
    assert len(ids) > 100
    V, I = ico_geometry()
    for j, gid in enumerate(ids[:60]):
        op, sc, q = act[gid, 0], act[gid, 1:4], act[gid, 4:8]
        s = math.sqrt(2 * math.log(op / alpha))
        R = quat_to_R(q)
        Sigma = R @ np.diag(sc ** 2) @ R.T
        M = o2w[j].astype(np.float64).reshape(4, 3).T  # 3x4 [R S s | mu]
        A, t = M[:, :3], M[:, 3]
        assert np.allclose(t, raw[gid, :3])
        Ainv_T = np.linalg.inv(A).T
        for f in range(20):
            p0, p1, p2 = V[I[f]]
            n = np.cross(p0 - p2, p1 - p0)
            nh = n / np.linalg.norm(n)
            a = Ainv_T @ nh
            dm = 1.0 / math.sqrt(a @ Sigma @ a)
            assert abs(dm / s - 1) < 2e-5, (gid, f, dm, s)
        Wm = w2o[j].astype(np.float64).reshape(4, 3).T
        comp = Wm[:, :3] @ np.vstack([A.T, t]).T[:, :3]
        assert np.allclose(comp, np.eye(3), atol=2e-5)
        assert np.allclose(Wm[:, :3] @ t + Wm[:, 3], 0, atol=2e-5 * (1 + np.abs(Wm[:, 3]).max()))


def check_tri_vs_mt():
    rng = np.random.default_rng(5)
    V, I = ico_geometry()
    agree = 0
    for _ in range(3000):
        o = rng.normal(size=3) * 3
        tgt = rng.normal(size=3) * 0.6
        d = tgt - o
        d /= np.linalg.norm(d)
        o32, d32 = o.astype(F32), d.astype(F32)
        for f in range(20):
            p0, p1, p2 = V[I[f]]
            t, u, v = mt64(o32.astype(np.float64), d32.astype(np.float64), p0, p1, p2)
            hit, t32 = orc.tri_hit(o32, d32, 1e-3, 1e5, f)
            margin = min(u, v, 1 - u - v)
            if abs(margin) < 1e-5 or abs(t - 1e-3) < 1e-5:
                continue
            expect = margin > 0 and t > 1e-3
            assert hit == expect, (o, d, f, t, u, v)
            if hit:
                assert abs(t32 - t) <= 1e-5 * max(1.0, abs(t))
                agree += 1
    assert agree > 500


def _fp64_first_round(raw, alpha, u, W, H, k):
    """All front-face proxy hits in (t_min, t_max) per ray by FP64
    Moller-Trumbore on the double-precision inverse of the proxy transform,
    sorted by t; returns per-ray (ids, ambiguous flag)."""
    packed, act = orc.activate(raw)
    ids, o2w, w2o = orc.build_proxies(raw, act, alpha)
    V, I = ico_geometry()
    p = orc.params_from_uniforms(u, W, H)
    o = np.array(u.to_world[12:15], np.float64)
    out = []
    for py in range(H):
        for px in range(W):
            d = orc.ray_dir(p, px, py).astype(np.float64)
            hits = []
            amb = False
            for j in range(len(ids)):
                A = o2w[j].astype(np.float64).reshape(4, 3).T
                Ainv = np.linalg.inv(A[:, :3])
                oo = Ainv @ (o - A[:, 3])
                od = Ainv @ d
                for f in range(20):
                    p0, p1, p2 = V[I[f]]
                    n = np.cross(p0 - p2, p1 - p0)
                    if n @ od >= 0:  # back face: n_w . d = n . (W2O d)
                        if abs(n @ od) < 1e-6:
                            amb = True
                        continue
                    t, uu, vv = mt64(oo, od, p0, p1, p2)
                    m = min(uu, vv, 1 - uu - vv)
                    if abs(m) < 1e-5 * (1 + abs(t)):
                        amb = True
                    if m < 0 or not (u.t_min < t < u.t_max):
                        continue
                    hits.append((t, int(ids[j])))
            hits.sort()
            ts = [h[0] for h in hits]
            if any(abs(a - b) < 1e-5 for a, b in zip(ts, ts[1:])):
                amb = True
            out.append(([h[1] for h in hits[:k]], amb))
    return out


def check_kbuffer_bruteforce():
    raw = gt.generate_raw(n=150, seed=12, config="c1", box_half=0.6, log_scale_mean=-2.0)
    alpha, W, H, k = 0.05, 12, 12, 4
    tw = gt.camera_look_at([0.0, 0.0, 3.0], [0.0, 0.0, 0.0])
    u = gt.make_uniforms(tw, W, H, hfov_deg=40, alpha_min=alpha, k=k, T_min=0.0)
    ref = _fp64_first_round(raw, alpha, u, W, H, k)
    s = orc.OracleScene(raw, alpha)
    r = orc.render(s, orc.params_from_uniforms(u, W, H), max_rounds=8, use_bvh=False)
    checked = hit_rays = 0
    for i, (ids, amb) in enumerate(ref):
        if amb:
            continue
        got = [x for x in r["hit_ids"][i, 0] if x != orc.SENTINEL]
        assert got == ids, (i, got, ids)
        checked += 1
        hit_rays += bool(ids)
    assert checked > 100 and hit_rays > 30


def check_response():
    rng = np.random.default_rng(7)
    # isotropic: alpha = exp(-b^2 / sigma^2) at impact parameter b (no 1/2)
    for _ in range(200):
        mu = rng.normal(size=3)
        ls = rng.uniform(-3, 0)
        q = rng.normal(size=4)
        rec = orc.activate(packed_record(mu, [ls] * 3, q, 2.0))[0][0]
        sig = float(rec[4])
        o = (mu + rng.normal(size=3) * 4).astype(F32)
        d = (mu + rng.normal(size=3) * sig * 0.8 - o)
        d = (d / np.linalg.norm(d)).astype(F32)
        m, od, dd = rec[:3].astype(np.float64), o.astype(np.float64), d.astype(np.float64)
        v = m - od
        b2 = v @ v - (v @ dd) ** 2 / (dd @ dd)
        ref = math.exp(-b2 / sig ** 2)
        got = orc.response(rec, o, d)
        assert abs(got - ref) <= 3e-5 * max(ref, 1e-3), (got, ref)
    # ray through the mean: response 1
    rec = orc.activate(packed_record([0.5, -0.2, 0.1], [-1.0, -2.0, -1.5], [0.3, 0.1, -0.7, 0.2], 1.0))[0][0]
    d = np.array([0.0, 0.0, -1.0], F32)
    o = (rec[:3] - 4 * d).astype(F32)
    assert abs(orc.response(rec, o, d) - 1.0) < 1e-5
    # anisotropic: exp(-min_t (o + t d - mu)^T Sigma^-1 (o + t d - mu)), Sigma = R S^2 R^T
    for _ in range(200):
        mu = rng.normal(size=3)
        q = rng.normal(size=4)
        rec = orc.activate(packed_record(mu, rng.uniform(-3, -0.5, 3), q, 2.0))[0][0]
        R = quat_to_R(rec[8:12].astype(np.float64))
        Sinv = R @ np.diag(1.0 / rec[4:7].astype(np.float64) ** 2) @ R.T
        o = (mu + rng.normal(size=3) * 3).astype(F32)
        d = (mu + rng.normal(size=3) * 0.3 - o)
        d = (d / np.linalg.norm(d)).astype(F32)
        od, dd, m = o.astype(np.float64), d.astype(np.float64), rec[:3].astype(np.float64)
        t = ((m - od) @ Sinv @ dd) / (dd @ Sinv @ dd)
        e = od + t * dd - m
        ref = math.exp(-(e @ Sinv @ e))
        got = orc.response(rec, o, d)
        assert abs(got - ref) <= 1e-4 * max(ref, 1e-2) + 1e-6, (got, ref)


def _real_sh(l, m, d):
    theta, phi = math.acos(max(-1.0, min(1.0, d[2]))), math.atan2(d[1], d[0])
    Y = lambda ll, mm: complex(sps.sph_harm_y(ll, mm, theta, phi))
    if m < 0:
        return math.sqrt(2) * Y(l, -m).imag
    if m == 0:
        return Y(l, 0).real
    return math.sqrt(2) * Y(l, m).real


def check_sh():
    """Radiance = max(0, 0.5 + sum_lm Y_lm(d) c_lm) with the real SH built from
    scipy's complex SH (Condon-Shortley phase), sh order (l, m=-l..l)."""
    rng = np.random.default_rng(8)
    for _ in range(60):
        sh = rng.normal(size=(16, 3)) * 0.4
        rec = np.zeros(80, F32)
        for i in range(16):
            rec[16 + 4 * i:16 + 4 * i + 3] = sh[i]
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(F32)
        dd = d.astype(np.float64)
        for deg in range(4):
            basis = [_real_sh(l, m, dd) for l in range(deg + 1) for m in range(-l, l + 1)]
            ref = np.maximum(0.0, 0.5 + np.array(basis) @ sh[:len(basis)].astype(F32).astype(np.float64))
            got = orc.radiance(rec, deg, d)
            assert np.abs(got - ref).max() < 2e-5, (deg, got, ref)


def _axis_scene(gaussians, k, T_min=1e-3, alpha=0.01):
    """Gaussians (mu, log_scale, quat, logit, dc) seen by a 1x1 camera at
    (0,0,10) looking down -z; returns (oracle result, activated records)."""
    raw = np.concatenate([packed_record(*g) for g in gaussians])
    s = orc.OracleScene(raw, alpha)
    tw = gt.camera_look_at([0.0, 0.0, 10.0], [0.0, 0.0, 0.0])
    u = gt.make_uniforms(tw, 1, 1, hfov_deg=30, alpha_min=alpha, k=k, T_min=T_min, sh_degree_max=0)
    r = orc.render(s, orc.params_from_uniforms(u, 1, 1), max_rounds=8, use_bvh=False)
    return r, s.packed


QA = [0.9, 0.3, -0.2, 0.25]   # generic orientations: the axis ray enters through a face interior
QB = [0.4, -0.5, 0.6, 0.1]


def check_composite_closed_form():
    ga = ([0, 0, 1.0], [-1.5] * 3, QA, 0.4, [0.8, -0.3, 0.1])
    gb = ([0, 0, -1.0], [-1.2] * 3, QB, 1.1, [-0.2, 0.6, 0.4])
    r, pk = _axis_scene([ga, gb], k=4)
    ids = [x for x in r["hit_ids"][0, 0] if x != orc.SENTINEL]
    assert ids == [0, 1], ids
    C0 = 0.28209479177387814
    oa, ob = float(pk[0, 12]), float(pk[1, 12])  # the axis ray passes through both means: response 1
    ca = np.maximum(0, 0.5 + C0 * pk[0, 16:19].astype(np.float64))
    cb = np.maximum(0, 0.5 + C0 * pk[1, 16:19].astype(np.float64))
    ref = oa * ca + (1 - oa) * ob * cb
    assert np.abs(r["rgba"][0, :3] - ref).max() < 1e-5, (r["rgba"][0], ref)
    assert abs(r["T"][0] - (1 - oa) * (1 - ob)) < 1e-6


def check_relaunch_window():
    """k=1: a second Gaussian whose proxy is the first one shifted by delta
    along the ray enters at t1 + delta.  t_curr = t1 + 1e-4 (GS_shader.ts:248)
    skips it for delta = 5e-5 and keeps it for delta = 3e-4."""
    C0 = 0.28209479177387814
    for delta, kept in [(5e-5, False), (3e-4, True)]:
        ga = ([0, 0, 0.0], [-1.5] * 3, QA, 0.0, [0.8, -0.3, 0.1])
        gb = ([0, 0, -delta], [-1.5] * 3, QA, 0.0, [-0.2, 0.6, 0.4])
        r, pk = _axis_scene([ga, gb], k=1)
        oa = float(pk[0, 12])
        ca = np.maximum(0, 0.5 + C0 * pk[0, 16:19].astype(np.float64))
        cb = np.maximum(0, 0.5 + C0 * pk[1, 16:19].astype(np.float64))
        ref = oa * ca + ((1 - oa) * oa * cb if kept else 0)
        assert np.abs(r["rgba"][0, :3] - ref).max() < 1e-5, (delta, r["rgba"][0], ref)
        assert r["n_rounds"][0] == (3 if kept else 2)


def check_invariants_and_reeval():
    raw = gt.generate_raw(n=30000, seed=13, config="c4")
    s = orc.OracleScene(raw, 0.02)
    tw = gt.camera_look_at([0.4, -0.3, 3.0], [0, 0, 0])
    W, H = 40, 32
    u = gt.make_uniforms(tw, W, H, hfov_deg=45, alpha_min=0.02, k=8, T_min=0.001)
    p = orc.params_from_uniforms(u, W, H)
    r = orc.render(s, p, max_rounds=64)
    assert r["stats"]["overflow"] == 0
    assert np.all(r["T"] > 0) and np.all(r["T"] <= 1)
    assert np.all(r["rgba"][:, :3] >= 0) and np.all(r["rgba"][:, 3] == 1)
    op = s.packed[:, 12]
    proxy = set(s.proxy_ids.tolist())
    for i in range(W * H):
        nr = r["n_rounds"][i]
        for rd in range(nr):
            row = r["hit_ids"][i, rd]
            sent = np.nonzero(row == orc.SENTINEL)[0]
            if sent.size:
                assert np.all(row[sent[0]:] == orc.SENTINEL) and rd == nr - 1
            for g in row[row != orc.SENTINEL]:
                assert int(g) in proxy and op[g] > 0.0
            acc = r["accept"][i, rd]
            assert acc >> int(sent[0] if sent.size else u.k) == 0
        assert np.all(r["hit_ids"][i, nr:] == orc.SENTINEL)
    assert r["stats"]["h_comp"] > 500 and (r["n_rounds"] > 1).sum() > 50
    rgb64, T64 = orc.reeval_f64(s, p, r["n_rounds"], r["hit_ids"], r["accept"])
    assert np.abs(r["rgba"][:, :3] - rgb64).max() < 1e-4
    assert np.abs(r["T"] - T64).max() < 1e-5


def _fp64_hits(raw, alpha, u, W, H):
    """Every proxy face the line of each camera ray crosses in (t_min, t_max),
    by FP64 Moller-Trumbore on the double inverse of each proxy transform:
    per ray (hits sorted by t as (t, Gaussian id, front), ambiguous flag).
    Front means n . (W2O d) < 0 with the outward normal n = cross(e1, e0)
    (the facing rule of trace.glsl:153-164 written with n_w . d = n . W2O d).
    A ray is ambiguous when a crossing lies within 1e-5 of a face edge, of
    t_min or of grazing incidence, or two crossings lie within 1e-5 in t."""
    packed, act = orc.activate(raw)
    ids, o2w, _ = orc.build_proxies(raw, act, alpha)
    V, I = ico_geometry()
    P0, P1, P2 = V[I[:, 0]], V[I[:, 1]], V[I[:, 2]]
    E1, E2 = P1 - P0, P2 - P0
    Nrm = np.cross(P0 - P2, P1 - P0)
    p = orc.params_from_uniforms(u, W, H)
    o = np.array(u.to_world[12:15], np.float64)
    A = o2w.astype(np.float64).reshape(-1, 4, 3).transpose(0, 2, 1)
    Ainv = np.linalg.inv(A[:, :, :3])
    oo = np.einsum("nij,nj->ni", Ainv, o[None, :] - A[:, :, 3])
    out = []
    for py in range(H):
        for px in range(W):
            d = orc.ray_dir(p, px, py).astype(np.float64)
            od = Ainv @ d
            with np.errstate(divide="ignore", invalid="ignore"):
                h = np.cross(od[:, None, :], E2[None, :, :])
                f = 1.0 / (E1[None] * h).sum(-1)
                sv = oo[:, None, :] - P0[None]
                uu = f * (sv * h).sum(-1)
                q = np.cross(sv, E1[None])
                vv = f * (od[:, None, :] * q).sum(-1)
                t = f * (E2[None] * q).sum(-1)
            m = np.minimum(np.minimum(uu, vv), 1 - uu - vv)
            ndot = (Nrm[None] * od[:, None, :]).sum(-1)
            ok = np.isfinite(t) & np.isfinite(m)
            near_edge = ok & (np.abs(m) < 1e-5 * (1 + np.abs(t))) & (t > u.t_min - 1e-4)
            inside = ok & (m > 0) & (t > u.t_min) & (t < u.t_max)
            amb = bool(near_edge.any()) or bool((inside & (np.abs(ndot) < 1e-6)).any()) or \
                bool((ok & (m > 0) & (np.abs(t - u.t_min) < 1e-5)).any())
            jj, ff = np.nonzero(inside)
            hits = sorted((float(t[j, k]), int(ids[j]), bool(ndot[j, k] < 0)) for j, k in zip(jj, ff))
            ts = [x[0] for x in hits]
            if any(abs(a - b) < 1e-5 * (1 + abs(a)) for a, b in zip(ts, ts[1:])):
                amb = True
            out.append((hits, amb))
    return out


def _radiance_real_sh(rec, deg, d):
    """max(0, 0.5 + sum_lm Y_lm(d) sh_lm) with scipy's real SH (the SH pin's basis)."""
    dd = np.asarray(d, np.float64)
    basis = [_real_sh(l, m, dd) for l in range(deg + 1) for m in range(-l, l + 1)]
    sh = np.array([rec[16 + 4 * i:16 + 4 * i + 3] for i in range(len(basis))], np.float64)
    return np.maximum(0.0, 0.5 + np.array(basis) @ sh)


def _debug_scenes():
    """A small random scene seen from outside and from inside its volume."""
    raw = gt.generate_raw(n=120, seed=14, config="c1", box_half=0.6, log_scale_mean=-1.8)
    return raw, [([0.0, 0.0, 3.0], [0.0, 0.0, 0.0]), ([0.05, -0.1, 0.15], [0.3, 0.2, -2.0])]


def check_debug1_back_face():
    """debugMode 1 traces with no cull flag (GS_shader.ts:292-294): a camera
    inside a proxy sees that Gaussian through its exit (back) face, which
    debugMode 0 culls; the colour is the SH radiance along the ray direction
    (GS_shader.ts:318-319), not along normalize(mu - o)."""
    ga = ([0.0, 0.0, 0.0], [-0.5] * 3, QA, 3.0, [0.8, -0.3, 0.1], np.linspace(-0.3, 0.3, 45))
    gb = ([0.0, 0.0, -6.0], [-1.0] * 3, QB, 2.0, [-0.2, 0.6, 0.4], np.linspace(0.2, -0.2, 45))
    raw = np.concatenate([packed_record(*ga), packed_record(*gb)])
    s = orc.OracleScene(raw, 0.01)
    tw = gt.camera_look_at([0.1, 0.2, 0.3], [0.4, -0.3, -5.0])
    for mode in (0, 1):
        u = gt.make_uniforms(tw, 1, 1, hfov_deg=30, alpha_min=0.01, k=4, T_min=1e-3, debug_mode=mode)
        p = orc.params_from_uniforms(u, 1, 1)
        r = orc.render(s, p, max_rounds=4, use_bvh=False)
        d = orc.ray_dir(p, 0, 0)
        first = int(r["hit_ids"][0, 0, 0])
        if mode == 0:
            assert first == 1, "mode 0: the enclosing proxy has no front face ahead; Gaussian 1 is the first hit"
        else:
            assert first == 0, "mode 1: the enclosing proxy's exit face is the closest hit"
            ref = _radiance_real_sh(s.packed[0], 3, d)
            assert np.abs(r["rgba"][0, :3] - ref).max() < 2e-5, (r["rgba"][0], ref)
            away = np.asarray(s.packed[0][:3], np.float64) - np.array(u.to_world[12:15], np.float64)
            other = _radiance_real_sh(s.packed[0], 3, away / np.linalg.norm(away))
            assert np.abs(ref - other).max() > 1e-3  # the two directions give distinguishable colours


def check_debug1_closest():
    """debugMode 1 = the closest proxy face in (t_min, t_max), front or back,
    coloured by radiance along the ray direction; black when there is none."""
    raw, cams = _debug_scenes()
    alpha, W, H = 0.05, 12, 10
    checked = back = empty = 0
    for pos, tgt in cams:
        tw = gt.camera_look_at(pos, tgt)
        u = gt.make_uniforms(tw, W, H, hfov_deg=50, alpha_min=alpha, k=6, debug_mode=1)
        p = orc.params_from_uniforms(u, W, H)
        ref = _fp64_hits(raw, alpha, u, W, H)
        s = orc.OracleScene(raw, alpha)
        r = orc.render(s, p, max_rounds=1, use_bvh=False)
        for i, (hits, amb) in enumerate(ref):
            if amb:
                continue
            got = int(r["hit_ids"][i, 0, 0])
            if not hits:
                assert got == orc.SENTINEL and np.all(r["rgba"][i, :3] == 0), i
                empty += 1
                continue
            assert got == hits[0][1], (i, got, hits[:2])
            back += not hits[0][2]
            col = _radiance_real_sh(s.packed[got], 3, orc.ray_dir(p, i % W, i // W))
            assert np.abs(r["rgba"][i, :3] - col).max() < 2e-5, (i, r["rgba"][i], col)
            checked += 1
    assert checked > 100 and back > 8 and empty >= 3, (checked, back, empty)


BINS = [(0, (0, 0, 0)), (20, (0.5, 0.5, 0)), (40, (0.5, 0, 0.5)), (60, (0, 0.5, 0.5)), (80, (0.5, 0, 0)),
        (100, (0, 0.5, 0)), (150, (0, 0, 0.5))]


def _bin_colour(c):
    """The hit-count colour table of GS_shader.ts:212-227: 0 is black, then
    bins [1,20) [20,40) [40,60) [60,80) [80,100) [100,150), then grey."""
    if c == 0:
        return (0, 0, 0)
    for hi, col in BINS[1:]:
        if c < hi:
            return col
    return (0.5, 0.5, 0.5)


def check_debug2_count():
    """debugMode 2's count = the front-face crossings in (t_min, t_max)
    (brute force in FP64); the colour is the table's bin of that count."""
    raw, cams = _debug_scenes()
    alpha, W, H = 0.05, 12, 10
    checked = multi = 0
    for pos, tgt in cams:
        tw = gt.camera_look_at(pos, tgt)
        u = gt.make_uniforms(tw, W, H, hfov_deg=50, alpha_min=alpha, k=6, debug_mode=2)
        p = orc.params_from_uniforms(u, W, H)
        ref = _fp64_hits(raw, alpha, u, W, H)
        r = orc.render(orc.OracleScene(raw, alpha), p, max_rounds=1, use_bvh=False)
        for i, (hits, amb) in enumerate(ref):
            if amb:
                continue
            c = sum(1 for h in hits if h[2])
            assert int(r["hit_ids"][i, 0, 0]) == c, (i, int(r["hit_ids"][i, 0, 0]), c)
            assert tuple(float(x) for x in r["rgba"][i, :3]) == _bin_colour(c)
            checked += 1
            multi += c >= 3
    assert checked > 100 and multi > 20, (checked, multi)


def _count_scene(c, seed):
    """c Gaussians on the camera axis whose proxies the 1x1 camera ray enters
    once each (orientations redrawn until FP64 brute force finds exactly c
    unambiguous front-face crossings)."""
    rng = np.random.default_rng(seed)
    tw = gt.camera_look_at([0.0, 0.0, 10.0], [0.0, 0.0, 0.0])
    u = gt.make_uniforms(tw, 1, 1, hfov_deg=30, alpha_min=0.01, k=6, debug_mode=2)
    zs = np.linspace(5.0, -8.0, max(c, 1))
    for _ in range(20):
        gs = [([0.0, 0.0, float(zs[j])], [-1.5] * 3, rng.normal(size=4), 0.4) for j in range(c)]
        raw = np.concatenate([packed_record(*g) for g in gs]) if c else \
            packed_record([0, 0, -50.0], [-1.5] * 3, [1, 0, 0, 0], -20.0)  # no proxy (opacity < alpha_min)
        hits, amb = _fp64_hits(raw, 0.01, u, 1, 1)[0]
        if not amb and sum(1 for h in hits if h[2]) == c:
            return raw, u
    raise AssertionError(f"no unambiguous scene for c={c}")


def check_debug2_bins():
    """Ray hit counts on each side of every bin boundary of the colour table."""
    for c in (0, 1, 19, 20, 39, 40, 59, 60, 79, 80, 99, 100, 149, 150, 151):
        raw, u = _count_scene(c, 100 + c)
        r = orc.render(orc.OracleScene(raw, 0.01), orc.params_from_uniforms(u, 1, 1), max_rounds=1, use_bvh=False)
        assert int(r["hit_ids"][0, 0, 0]) == c
        assert tuple(float(x) for x in r["rgba"][0, :3]) == _bin_colour(c), (c, r["rgba"][0])


def _anyhit_calls(ts, k, t_max):
    """Any-hit invocations of one debugMode-2 trace for hits arriving in the
    order ts (GS_shader.ts:201-210, 323-366; trace.glsl:183-215): a candidate
    reaches the any-hit shader only below the current tmax; the shader
    inserts it into k slots reset to t_max and commits (tmax = t) unless it
    lies strictly below the k-th slot after insertion."""
    kc = [t_max] * k
    tmax, calls = t_max, 0
    for t in ts:
        if not t < tmax:
            continue
        calls += 1
        h = t
        for i in range(k):
            if h < kc[i]:
                kc[i], h = h, kc[i]
        if not t < kc[k - 1]:
            tmax = t
    return calls


def check_debug2_reference_range():
    """The reference's debugMode-2 count depends on the arrival order of hits
    (its BVH traversal order): over all orders it spans [min(k, n), n] for n
    front-face crossings, reaching n only when they arrive far to near.  The
    oracle (and kernel) count n, the upper end, for every order."""
    raw, cams = _debug_scenes()
    alpha, W, H, k = 0.05, 12, 10, 6
    rng = np.random.default_rng(3)
    pos, tgt = cams[0]
    tw = gt.camera_look_at(pos, tgt)
    u = gt.make_uniforms(tw, W, H, hfov_deg=50, alpha_min=alpha, k=k, debug_mode=2)
    ref = _fp64_hits(raw, alpha, u, W, H)
    r = orc.render(orc.OracleScene(raw, alpha), orc.params_from_uniforms(u, W, H), max_rounds=1, use_bvh=False)
    seen = 0
    for i, (hits, amb) in enumerate(ref):
        ts = [h[0] for h in hits if h[2]]
        if amb or len(ts) <= k:
            continue
        calls = [_anyhit_calls(list(rng.permutation(ts)), k, u.t_max) for _ in range(60)]
        assert min(calls) >= k and max(calls) <= len(ts)
        assert _anyhit_calls(sorted(ts), k, u.t_max) == k
        assert _anyhit_calls(sorted(ts, reverse=True), k, u.t_max) == len(ts)
        assert int(r["hit_ids"][i, 0, 0]) == len(ts)
        seen += 1
    assert seen > 10


PINS = {
    "exp64": check_exp64, "log64": check_log64, "icosahedron": check_icosahedron, "ray_dir": check_ray_dir, "aa_jitter": check_aa_jitter,
    "proxy_tangent": check_proxy_tangent, "proxy_tangent_packed": lambda: check_proxy_tangent(packed_input=True),
    "tri_vs_mt": check_tri_vs_mt,
    "kbuffer_bruteforce": check_kbuffer_bruteforce, "response": check_response, "sh": check_sh,
    "composite": check_composite_closed_form, "relaunch": check_relaunch_window,
    "invariants": check_invariants_and_reeval,
    "debug1_back_face": check_debug1_back_face, "debug1_closest": check_debug1_closest,
    "debug2_count": check_debug2_count, "debug2_bins": check_debug2_bins,
    "debug2_reference_range": check_debug2_reference_range,
}


@pytest.mark.parametrize("name", list(PINS))
def test_pin(name):
    orc.set_mutation(0)
    PINS[name]()


# mutation -> the pin that must catch it
MUTATIONS = [
    (1, "response"),          # 1/2 factor in the exponent (GS_shader.ts:135 has none)
    (2, "sh"),                # sign of the degree-1 y term flipped
    (3, "ray_dir"),           # camera X axis not negated (GS_shader.ts:277)
    (4, "kbuffer_bruteforce"),  # facing test inverted (trace.glsl:159)
    (5, "relaunch"),          # relaunch without the 1e-4 step (GS_shader.ts:248)
    (6, "composite"),         # transmittance dropped from L += T*alpha*c
    (7, "proxy_tangent"),     # proxy scale s = sqrt(2 ln(o/alpha)) dropped (main.ts:84)
    (7, "proxy_tangent_packed"),
    (8, "debug1_back_face"),  # debugMode 1 culling back faces (GS_shader.ts:292-294 passes no cull flag)
    (8, "debug1_closest"),
    (9, "debug1_closest"),    # debugMode 1 coloured along normalize(mu - o) instead of the ray (GS_shader.ts:318-320)
    (9, "debug1_back_face"),
    (10, "debug2_bins"),      # hit-count bin threshold off by one (c <= 20, GS_shader.ts:214)
    (11, "debug2_count"),     # debugMode 2 counting back faces too (its trace culls them, GS_shader.ts:206-207)
]


@pytest.mark.parametrize("mutation,pin", MUTATIONS)
def test_mutation_detected(mutation, pin):
    orc.set_mutation(mutation)
    try:
        with pytest.raises(AssertionError):
            PINS[pin]()
    finally:
        orc.set_mutation(0)


def test_packed_buffer_proxies_vs_js_doubles():
    """How the proxies of a packed buffer (gt_scene_create_packed) differ from
    those of the JS doubles (gt_scene_create_raw, main.ts:73-94): the same
    Gaussians get proxies; each transform column agrees to 1e-4 of its norm
    (the inputs are rounded to f32 once more, and the proxy scale
    sqrt(2 ln(opacity / alpha_min)) amplifies the opacity's rounding for
    opacities near alpha_min: 1.3e-5 at worst on C1, 4e-7 on pc_short.ply);
    hit lists agree on every ray of
    pc_short.ply and on all but a handful of the 65,536 C1 rays, where a ray
    grazes a proxy face."""
    for name, raw, alpha, W, H, cam, k, Tmin, max_diff_rays in [
            ("pc_short", gt.read_ply(os.path.join(os.path.dirname(__file__), "golden", "pc_short.ply"))[0],
             0.01, 64, 64, None, 4, 1e-3, 0),
            ("c1", gt.generate_raw(seed=0, config="c1"), 0.1, 256, 256, ([0, 0, 8.0], [0, 0, 0.0]), 6, 0.03, 16)]:
        a = orc.OracleScene(raw, alpha)
        b = orc.OracleScene(packed=a.packed, alpha_min=alpha)
        assert np.array_equal(a.proxy_ids, b.proxy_ids), name
        for m1, m2 in ((a.o2w, b.o2w), (a.w2o, b.w2o)):
            c1 = m1.astype(np.float64).reshape(-1, 4, 3)[:, :3]
            c2 = m2.astype(np.float64).reshape(-1, 4, 3)[:, :3]
            rel = np.abs(c1 - c2).max(axis=2) / np.linalg.norm(c1, axis=2)
            assert rel.max() < 1e-4, (name, rel.max())
        if cam is None:
            c = raw[:, :3].mean(axis=0)
            cam = (list(c + np.array([0.0, 0.0, 3.0])), list(c))
        u = gt.make_uniforms(gt.camera_look_at(*cam), W, H, hfov_deg=60, alpha_min=alpha, k=k, T_min=Tmin)
        p = orc.params_from_uniforms(u, W, H)
        ra, rb = orc.render(a, p, max_rounds=24), orc.render(b, p, max_rounds=24)
        differ = int((ra["hit_ids"] != rb["hit_ids"]).any(axis=(1, 2)).sum())
        assert differ <= max_diff_rays, (name, differ)
