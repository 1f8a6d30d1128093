"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer,
and its threaded stages under ThreadSanitizer, without a GPU.

tests/sanitize/gt_sanitize_main.cpp drives the PLY reader (gt_util.cpp), the
host scene builder (gt_host_build.cpp: activations, proxies, binned-SAH tree,
4-wide collapse, shading records), the PNG writer (gt_png.cpp) and the
oracle (oracle/gt_oracle.c).  This module builds it with g++ / gcc only (no
hipcc, no device code) into a temporary directory and feeds it malformed,
truncated and fuzzed .ply files, records at the edges of f32, generated
scenes and PNG edge cases.  Every sanitizer stops the driver at its first
report (-fno-sanitize-recover, halt_on_error), which fails the test.
"""
import os
import shutil
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "webgpu-gaussiantracer_amd", "csrc")
CXX_SOURCES = [os.path.join(CSRC, f) for f in ("gt_host_build.cpp", "gt_util.cpp", "gt_png.cpp")] + [
    os.path.join(ROOT, "tests", "sanitize", "gt_sanitize_main.cpp")]
C_SOURCES = [os.path.join(ROOT, "oracle", "gt_oracle.c")]

SANITIZERS = {
    "asan": ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all"],
    "tsan": ["-fsanitize=thread"],
}
ENV = {
    "ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=1",
    "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
    "TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1",
}

REQUIRED = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
            "rot_0", "rot_1", "rot_2", "rot_3"]


def _build(out_dir, kind):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("g++/gcc not found")
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread", "-Wall"] + SANITIZERS[kind]
    jobs, objs = [], []
    for i, src in enumerate(CXX_SOURCES + C_SOURCES):
        obj = os.path.join(out_dir, f"{kind}_{i}.o")
        objs.append(obj)
        if src.endswith(".c"):
            cmd = ["gcc", "-std=gnu99", "-Wno-unused-function"] + flags + ["-c", src, "-o", obj]
        else:
            cmd = ["g++", "-std=c++17"] + flags + ["-c", src, "-o", obj]
        jobs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    for j in jobs:
        out = j.communicate()[0]
        assert j.returncode == 0, out
    exe = os.path.join(out_dir, f"gt_sanitize_{kind}")
    r = subprocess.run(["g++"] + flags + ["-o", exe] + objs + ["-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


@pytest.fixture(scope="module")
def asan(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("asan")), "asan")


@pytest.fixture(scope="module")
def tsan(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("tsan")), "tsan")


def _run(exe, *args, timeout=300):
    env = dict(os.environ, **ENV)
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout, env=env)
    report = r.stderr
    assert "Sanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], report[-4000:])
    assert r.stdout.strip().endswith("failures 0"), r.stdout[-2000:]
    return r.stdout


def _records(n, props, seed=0):
    """Plausible values for each named property (log-scales near -3, logits
    near 0, unit-ish quaternions, small SH)."""
    rng = np.random.default_rng(seed)
    cols = []
    for name in props:
        if name in ("x", "y", "z"):
            cols.append(rng.uniform(-2, 2, n))
        elif name.startswith("scale_"):
            cols.append(rng.normal(-3.0, 0.5, n))
        elif name == "opacity":
            cols.append(rng.normal(0.0, 1.5, n))
        elif name.startswith("rot_"):
            cols.append(rng.normal(0.0, 1.0, n))
        elif name.startswith("f_dc_"):
            cols.append(rng.normal(0.0, 0.5, n))
        else:
            cols.append(rng.normal(0.0, 0.05, n))
    return np.stack(cols, axis=1) if cols else np.zeros((n, 0))


def _ply_bytes(props, n, types=None, values=None, fmt=b"binary_little_endian", count_text=None,
               extra_header=b"", eol=b"\n", trailing=b""):
    """A .ply file: props named, with PLY types (default float)."""
    types = types or {}
    vals = _records(n, props) if values is None else values
    hdr = [b"ply", b"format " + fmt + b" 1.0", b"comment made by the sanitizer test"]
    if extra_header:
        hdr.append(extra_header.rstrip(b"\n"))
    hdr.append(b"element vertex " + (count_text if count_text is not None else str(n).encode()))
    for p in props:
        hdr.append(b"property " + types.get(p, "float").encode() + b" " + p.encode())
    hdr.append(b"end_header")
    codes = {"float": "<f", "double": "<d", "uchar": "<B", "char": "<b", "int": "<i", "uint": "<I",
             "short": "<h", "ushort": "<H"}
    body = bytearray()
    for v in range(n):
        for j, p in enumerate(props):
            t = types.get(p, "float")
            t = t if t in codes else "float"  # a type the reader refuses: its header line is the test
            x = vals[v, j]
            if t in ("uchar", "char", "int", "uint", "short", "ushort"):
                x = int(np.clip(x, 0, 255))
            body += struct.pack(codes[t], x)
    return eol.join(hdr) + eol + bytes(body) + trailing


def _deg_props(deg):
    nrest = 3 * ((deg + 1) ** 2 - 1)
    return REQUIRED[:6] + ["nx", "ny", "nz"] + [f"f_rest_{i}" for i in range(nrest)] + REQUIRED[6:]


def _parse(out):
    res = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 5 and parts[0].endswith(".ply"):
            res[os.path.basename(parts[0])] = tuple(int(x) for x in parts[1:])
    return res


def test_sanitized_ply_corpus(asan, tmp_path):
    """Valid files of SH degree 0, 1 and 3 (with ignored normals, uchar and
    double properties, CRLF line ends, trailing bytes, an empty element
    before the vertices); every truncation of one file; 300 random edits;
    the malformed headers; records that describe no finite Gaussian."""
    files, expect = [], {}

    def add(name, data, rc=None, n=None, deg=None, brc=None):
        p = tmp_path / name
        p.write_bytes(data)
        files.append(str(p))
        expect[name] = (rc, n, deg, brc)

    add("deg0.ply", _ply_bytes(REQUIRED, 50), 0, 50, 0, 0)
    add("deg1.ply", _ply_bytes(_deg_props(1), 40), 0, 40, 1, 0)
    deg3 = _ply_bytes(_deg_props(3), 30)
    add("deg3.ply", deg3, 0, 30, 3, 0)
    add("types.ply", _ply_bytes(_deg_props(1) + ["flag", "weight"], 20, types={"flag": "uchar", "weight": "double",
                                                                                "nx": "double"}), 0, 20, 1, 0)
    add("crlf.ply", _ply_bytes(REQUIRED, 10, eol=b"\r\n"), 0, 10, 0, 0)
    add("trailing.ply", _ply_bytes(REQUIRED, 10, trailing=b"\x00garbage" * 9), 0, 10, 0, 0)
    add("face0.ply", _ply_bytes(REQUIRED, 10, extra_header=b"element face 0\nproperty list uchar int vertex_indices"),
        0, 10, 0, 0)
    add("empty_vertex.ply", _ply_bytes(REQUIRED, 0), 0, 0, 0, 0)

    # malformed headers and files
    add("empty.ply", b"", -4)
    add("noply.ply", b"plx\n" + _ply_bytes(REQUIRED, 2)[4:], -4)
    add("ascii.ply", _ply_bytes(REQUIRED, 2, fmt=b"ascii"), -4)
    add("bigendian.ply", _ply_bytes(REQUIRED, 2, fmt=b"binary_big_endian"), -4)
    add("huge_count.ply", _ply_bytes(REQUIRED, 2, count_text=b"4294967295"), -4)
    add("too_many.ply", _ply_bytes(REQUIRED, 2, count_text=b"99999999999999"), -5)
    for i, c in enumerate([b"-3", b"abc", b"", b"1e9", b"18446744073709551615", b"99999999999999999999"]):
        add(f"count{i}.ply", _ply_bytes(REQUIRED, 2, count_text=c), -4)
    add("rest46.ply", _ply_bytes(REQUIRED[:6] + [f"f_rest_{i}" for i in range(46)] + REQUIRED[6:], 2), -4)
    add("rest_deg4.ply", _ply_bytes(REQUIRED[:6] + [f"f_rest_{i}" for i in range(72)] + REQUIRED[6:], 2), -4)
    for bad in ["f_dc_3", "rot_4", "scale_3", "f_dc_-1", "rot_0001234567", "f_rest_", "scale_"]:
        add(f"idx_{bad}.ply", _ply_bytes(REQUIRED + [bad], 2), -4)
    add("missing.ply", _ply_bytes([x for x in REQUIRED if x != "rot_2"], 2), -4)
    add("dup.ply", _ply_bytes(REQUIRED + ["y"], 2), -4)
    add("face_first.ply", _ply_bytes(REQUIRED, 2, extra_header=b"element face 3\nproperty uchar n"), -4)
    add("two_vertex.ply", _ply_bytes(REQUIRED, 2, extra_header=b"element vertex 1\nproperty float x"), -4)
    add("list_prop.ply", _ply_bytes(REQUIRED, 2, types={"x": "list"}), -4)
    add("no_end.ply", _ply_bytes(REQUIRED, 2).replace(b"end_header", b"end_headr"), -4)
    add("long_header.ply", b"ply\nformat binary_little_endian 1.0\n" + b"comment x\n" * 100001 +
        _ply_bytes(REQUIRED, 1)[len(b"ply\nformat binary_little_endian 1.0\n"):], -4)

    # records at the edges: the reader accepts them; the builder refuses them
    names = REQUIRED
    base = _records(8, names, seed=5)
    col = {nm: j for j, nm in enumerate(names)}
    base[:, col["opacity"]] = 3.0  # opacity 0.95: every record has a proxy
    for label, field, val in [("pos_max", "x", 3.0e38), ("pos_nan", "y", np.nan), ("pos_inf", "z", -np.inf),
                              ("scale_huge", "scale_1", 88.5), ("scale_tiny", "scale_0", -87.0),
                              ("scale_overflow", "scale_2", 200.0), ("dc_nan", "f_dc_1", np.nan)]:
        v = base.copy()
        v[6, col[field]] = val
        add(f"edge_{label}.ply", _ply_bytes(names, 8, values=v), 0, 8, 0, -1)
    v = base.copy()
    v[6, col["rot_0"]:col["rot_3"] + 1] = 0.0
    add("edge_zero_quat.ply", _ply_bytes(names, 8, values=v), 0, 8, 0, -1)
    v = base.copy()
    v[6, col["x"]] = 1.0e30  # finite after padding: still a Gaussian, far away
    add("edge_far.ply", _ply_bytes(names, 8, values=v), 0, 8, 0, 0)
    v = base.copy()
    v[2, col["z"]] = 1e300  # a double property beyond f32: inf after the reader's conversion
    add("edge_double_overflow.ply", _ply_bytes(names, 8, values=v, types={"z": "double"}), 0, 8, 0, -1)

    # every 11th truncation of a valid degree-3 file
    for cut in range(0, len(deg3) - 1, 11):
        add(f"cut{cut}.ply", deg3[:cut], -4)
    add("cut_last.ply", deg3[:-1], -4)

    # random edits of the degree-1 file
    src = _ply_bytes(_deg_props(1), 7)
    rng = np.random.default_rng(0)
    for i in range(300):
        b = bytearray(src)
        for _ in range(int(rng.integers(1, 4))):
            pos = int(rng.integers(0, len(b)))
            op = int(rng.integers(0, 3))
            if op == 0:
                b[pos] = int(rng.integers(0, 256))
            elif op == 1:
                del b[pos:pos + int(rng.integers(1, 8))]
            else:
                b[pos:pos] = bytes(rng.integers(0, 256, size=int(rng.integers(1, 8))).astype(np.uint8))
        add(f"fuzz{i}.ply", bytes(b), None)

    files.append(str(tmp_path / "does_not_exist.ply"))
    expect["does_not_exist.ply"] = (-4, None, None, None)
    res = _parse(_run(asan, "ply", *files))
    assert len(res) == len(expect)
    outcomes = {}
    for name, (rc, n, deg, brc) in expect.items():
        got = res[name]
        if rc is not None:
            assert got[0] == rc, (name, got)
        else:
            assert got[0] in (0, -3, -4, -5), (name, got)
        if n is not None:
            assert got[1:] == (n, deg, brc), (name, got)
        if got[0] == 0:
            assert got[3] in (0, -1), (name, got)
        outcomes[got[0]] = outcomes.get(got[0], 0) + 1
    assert outcomes.get(0, 0) > 20 and outcomes.get(-4, 0) > 100


@pytest.mark.parametrize("n,seed,clusters,alpha,render", [
    (0, 1, 0, 0.01, 1),       # empty scene: no tree, every ray misses
    (1, 1, 0, 0.01, 1),       # one proxy: a root with one leaf
    (2000, 1, 0, 0.01, 1),    # uniform box
    (1500, 2, 4, 0.05, 1),    # clustered
    (300, 3, 1, 0.001, 1),    # one dense cluster, small alpha_min: deep hit lists
    (70000, 4, 8, 0.01, 0),   # large enough for the threaded build stages
])
def test_sanitized_scene_build(asan, tmp_path, n, seed, clusters, alpha, render):
    """The host builder's tree checks (every proxy in one leaf, boxes nested,
    stack need and depth as recorded, records in leaf order, bit-equal to the
    oracle's activations and proxies) and, for the small scenes, the
    oracle's BVH and brute-force renders bit for bit and within 1e-4 of its
    FP64 re-evaluation."""
    png = tmp_path / "frame.png"
    out = _run(asan, "scene", n, seed, clusters, alpha, render, png)
    assert f"scene n={n} " in out
    if render:
        assert png.exists() and png.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"


def _png_pixels(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        ln, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        crc, = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])
        assert crc == zlib.crc32(typ + body) & 0xffffffff
        chunks[typ] = chunks.get(typ, b"") + body
        pos += 12 + ln
    w, h = struct.unpack(">II", chunks[b"IHDR"][:8])
    raw = zlib.decompress(chunks[b"IDAT"])
    assert len(raw) == h * (4 * w + 1)
    rows = [raw[y * (4 * w + 1):(y + 1) * (4 * w + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows)
    return w, h, b"".join(r[1:] for r in rows)


def test_sanitized_png(asan, tmp_path):
    _run(asan, "png", tmp_path)
    for name, (w, h) in {"e.png": (1, 1), "f.png": (4096, 3), "g.png": (3, 4096)}.items():
        pw, ph, px = _png_pixels(tmp_path / name)
        assert (pw, ph) == (w, h) and px == b"\x7f" * (4 * w * h)
    assert not (tmp_path / "b.png").exists() and not (tmp_path / "c.png").exists()


@pytest.mark.parametrize("args", [(70000, 4, 8, 0.01, 0), (3000, 5, 2, 0.01, 1)])
def test_thread_sanitized_build_and_oracle(tsan, tmp_path, args):
    """The threaded stages (activation and proxy chunks, the tree build's
    parallel bounds, binning and subtree workers) and the oracle's threaded
    render under ThreadSanitizer."""
    _run(tsan, "scene", *args, tmp_path / "t.png")
