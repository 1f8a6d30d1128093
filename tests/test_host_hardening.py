"""Untrusted inputs on the host, without a GPU: malformed .ply files and
scene records that describe no finite Gaussian.

The reference has no such checks: its loader reads property types it does
not know without advancing (src/GS_ply.ts:120-129) and its activations
divide by the quaternion norm (src/GS_ply.ts:229-233), so a zero quaternion
or a NaN parameter reaches the TLAS build as NaN.  Here each is an error
code with a message naming the problem, and no exception or abort crosses
the C-ABI.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import gtrace as gt

REQUIRED = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
            "rot_0", "rot_1", "rot_2", "rot_3"]


def _ply(path, names, n, body=None, count_text=None, extra_header=b""):
    hdr = b"ply\nformat binary_little_endian 1.0\n" + extra_header
    hdr += b"element vertex %s\n" % (count_text if count_text is not None else str(n).encode())
    hdr += b"".join(b"property float %s\n" % x.encode() for x in names) + b"end_header\n"
    if body is None:
        body = np.arange(n * len(names), dtype="<f4").tobytes()
    path.write_bytes(hdr + body)
    return str(path)


def _read_error(path):
    with pytest.raises(gt.GTError) as e:
        gt.read_ply(path)
    return e.value.code, str(e.value)


def test_ply_huge_vertex_count_is_io_error(tmp_path):
    """A small file whose header claims 2^32 - 1 vertices is refused before
    anything is sized by that count (no bad_alloc across the C-ABI)."""
    p = _ply(tmp_path / "huge.ply", REQUIRED, 2, count_text=b"4294967295")
    code, msg = _read_error(p)
    assert code == -4 and "truncated" in msg
    p = _ply(tmp_path / "huge2.ply", REQUIRED, 2, count_text=b"99999999999999")
    code, _ = _read_error(p)
    assert code == -5  # beyond 2^32 - 1 vertices: a documented limit


@pytest.mark.parametrize("count", [b"-3", b"abc", b"", b"1e9"])
def test_ply_bad_vertex_count(tmp_path, count):
    p = _ply(tmp_path / "cnt.ply", REQUIRED, 1, count_text=count)
    code, _ = _read_error(p)
    assert code == -4


def test_ply_f_rest_not_multiple_of_three(tmp_path):
    """46 f_rest properties: previously read as degree 3 with f_rest_45
    mapped onto the opacity column."""
    names = REQUIRED[:6] + ["opacity"] + [f"f_rest_{i}" for i in range(46)] + REQUIRED[7:]
    p = _ply(tmp_path / "rest46.ply", names, 3)
    code, msg = _read_error(p)
    assert code == -4 and "multiple of 3" in msg


@pytest.mark.parametrize("bad", ["f_dc_3", "rot_4", "scale_3", "f_rest_x", "f_dc_-1", "rot_0001234567"])
def test_ply_property_index_out_of_range(tmp_path, bad):
    extra = [f"f_rest_{i}" for i in range(8)] if bad.startswith("f_rest") else []  # 9 f_rest: degree 1
    p = _ply(tmp_path / "idx.ply", REQUIRED + extra + [bad], 2)
    code, msg = _read_error(p)
    assert code == -4 and bad in msg


def test_ply_missing_and_duplicate_properties(tmp_path):
    code, msg = _read_error(_ply(tmp_path / "miss.ply", [x for x in REQUIRED if x != "opacity"], 2))
    assert code == -4 and "required" in msg
    code, msg = _read_error(_ply(tmp_path / "dup.ply", REQUIRED + ["x"], 2))
    assert code == -4 and "twice" in msg


def test_ply_element_before_vertex(tmp_path):
    p = _ply(tmp_path / "face.ply", REQUIRED, 1, extra_header=b"element face 2\nproperty uchar n\n")
    code, msg = _read_error(p)
    assert code == -4 and "before the vertex" in msg
    # an empty element before the vertices has no records and is accepted
    p = _ply(tmp_path / "face0.ply", REQUIRED, 1, extra_header=b"element face 0\n")
    raw, deg = gt.read_ply(p)
    assert raw.shape == (1, 59) and deg == 0


def test_ply_truncated_and_exact(tmp_path):
    n = 5
    full = np.arange(n * len(REQUIRED), dtype="<f4").tobytes()
    code, _ = _read_error(_ply(tmp_path / "t.ply", REQUIRED, n, body=full[:-1]))
    assert code == -4
    raw, deg = gt.read_ply(_ply(tmp_path / "ok.ply", REQUIRED, n, body=full + b"trailing"))
    vals = np.frombuffer(full, "<f4").reshape(n, len(REQUIRED))
    cols = [0, 1, 2, 3, 4, 5, 51, 52, 53, 54, 55, 56, 57, 58]
    assert deg == 0 and np.array_equal(raw[:, cols], vals) and np.all(raw[:, 6:51] == 0)


def test_ply_fuzzed_headers_never_crash(tmp_path):
    """Random edits of a valid file (header and body) give either records or
    an error code, never a crash."""
    names = REQUIRED[:6] + ["opacity"] + [f"f_rest_{i}" for i in range(9)] + REQUIRED[7:]
    base = open(_ply(tmp_path / "base.ply", names, 7), "rb").read()
    rng = np.random.default_rng(0)
    outcomes = {"ok": 0, "err": 0}
    for i in range(400):
        b = bytearray(base)
        for _ in range(rng.integers(1, 4)):
            pos = int(rng.integers(0, len(b)))
            op = rng.integers(0, 3)
            if op == 0:
                b[pos] = int(rng.integers(0, 256))
            elif op == 1:
                del b[pos:pos + int(rng.integers(1, 8))]
            else:
                b[pos:pos] = bytes(rng.integers(0, 256, size=int(rng.integers(1, 8))).astype(np.uint8))
        p = tmp_path / f"f{i}.ply"
        p.write_bytes(bytes(b))
        try:
            raw, deg = gt.read_ply(str(p))
            assert raw.shape[1] == 59 and 0 <= deg <= 3
            outcomes["ok"] += 1
        except gt.GTError as e:
            assert e.code in (-4, -5, -3)
            outcomes["err"] += 1
    assert outcomes["ok"] > 0 and outcomes["err"] > 0


def _invalid_raws():
    base = gt.generate_raw(n=64, seed=3, config="c1")
    cases = []
    for label, col, val in [("nan mean", 1, np.nan), ("inf dc", 4, np.inf), ("nan rest", 30, np.nan),
                            ("inf logit", 51, -np.inf), ("nan log-scale", 53, np.nan),
                            ("scale overflows f32", 52, 100.0), ("scale underflows f32", 54, -200.0),
                            ("nan quaternion", 57, np.nan)]:
        r = base.copy()
        r[37, col] = val
        cases.append((label, r))
    r = base.copy()
    r[37, 55:59] = 0.0
    cases.append(("zero quaternion", r))
    return cases


@pytest.mark.parametrize("label,raw", _invalid_raws(), ids=[c[0] for c in _invalid_raws()])
def test_invalid_records_rejected(label, raw):
    """gt_host_prepare and gt_scene_create_raw refuse the record (index 37)
    with GT_E_INVALID before any device work."""
    with pytest.raises(gt.GTError) as e:
        gt.host_prepare(raw, 0.05)
    assert e.value.code == -1 and "Gaussian 37" in str(e.value)
    lib = gt.lib()
    h = C.c_void_p()
    rc = lib.gt_scene_create_raw(raw.ctypes.data_as(C.c_void_p), raw.shape[0], 0.05, C.byref(h))
    assert rc == -1 and "Gaussian 37" in gt.last_error() and not h


def _overflow_raws():
    """Finite records whose proxy is not finite in f32 (record 37)."""
    base = gt.generate_raw(n=64, seed=3, config="c1")
    base[37, 51] = 3.0  # opacity 0.95: the record has a proxy at alpha_min 0.05
    cases = []
    for label, col, val in [("position near FLT_MAX", 0, 3.0e38), ("scale x proxy factor overflows", 53, 88.5),
                            ("inverse overflows", 52, -87.0)]:
        r = base.copy()
        r[37, col] = val
        cases.append((label, r))
    return cases


@pytest.mark.parametrize("label,raw", _overflow_raws(), ids=[c[0] for c in _overflow_raws()])
def test_proxy_overflow_rejected(label, raw):
    """A record that gt_prep_valid accepts but whose proxy transform, inverse
    or box (centre) overflows f32 is GT_E_INVALID naming it; the same record
    below alpha_min (no proxy) is accepted."""
    with pytest.raises(gt.GTError) as e:
        gt.host_prepare(raw, 0.05)
    assert e.value.code == -1 and "Gaussian 37" in str(e.value) and "proxy" in str(e.value)
    r = raw.copy()
    r[37, 51] = -10.0  # opacity 4.5e-5 < alpha_min: no proxy, nothing to overflow
    h = gt.host_prepare(r, 0.05)
    assert 37 not in h["proxy_ids"].tolist()


def test_valid_extremes_accepted():
    """Finite records at the edges stay valid: a tiny but nonzero quaternion,
    log-scales whose exp is a normal f32, logits that saturate the sigmoid."""
    raw = gt.generate_raw(n=8, seed=4, config="c1")
    raw[0, 55:59] = [1e-30, 0, 0, 0]
    raw[1, 52:55] = [80.0, -80.0, 0.0]
    raw[2, 51] = -800.0  # opacity 0: no proxy
    raw[3, 51] = 800.0   # opacity 1
    h = gt.host_prepare(raw, 0.05)
    assert np.all(np.isfinite(h["packed"]))
    assert 2 not in h["proxy_ids"].tolist() and 3 in h["proxy_ids"].tolist()


def test_packed_entry_argument_errors_without_gpu():
    lib = gt.lib()
    h = C.c_void_p()
    raw = gt.generate_raw(n=16, seed=5, config="c1")
    packed = np.ascontiguousarray(gt.host_prepare(raw, 0.05)["packed"])
    P = packed.ctypes.data_as(C.c_void_p)
    assert lib.gt_scene_create_packed(None, 16, 0.05, C.byref(h)) == -1
    assert lib.gt_scene_create_packed(P, 16, 0.05, None) == -1
    for bad in (0.0, 1.0, float("nan")):
        assert lib.gt_scene_create_packed(P, 16, bad, C.byref(h)) == -1 and "alpha_min" in gt.last_error()
    for label, field, val in [("nan pos", 1, np.nan), ("zero scale", 5, 0.0), ("negative scale", 6, -1.0),
                              ("inf opacity", 12, np.inf), ("nan sh", 16 + 4 * 9 + 2, np.nan)]:
        p2 = packed.copy()
        p2[9, field] = val
        rc = lib.gt_scene_create_packed(p2.ctypes.data_as(C.c_void_p), 16, 0.05, C.byref(h))
        assert rc == -1 and "Gaussian 9" in gt.last_error(), label
    p2 = packed.copy()
    p2[11, 8:12] = 0.0
    rc = lib.gt_scene_create_packed(p2.ctypes.data_as(C.c_void_p), 16, 0.05, C.byref(h))
    assert rc == -1 and "Gaussian 11" in gt.last_error()
