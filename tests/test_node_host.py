"""The JavaScript host (webgpu-gaussiantracer_amd/node) over the Node-API
addon: host-only entry points compared with the Python binding of the same
library on the CPU; a render through the addon compared bit for bit with the
Python binding's render on the GPU."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import gtrace as gt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "webgpu-gaussiantracer_amd", "node")
ADDON = os.path.join(ROOT, "webgpu-gaussiantracer_amd", "lib", "gtrace.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is absent")


def _node(script, tmp_path):
    f = tmp_path / "s.js"
    f.write_text(f"const gt = require({json.dumps(NODE_DIR)});\nconst fs = require('fs');\n" + script)
    r = subprocess.run(["node", str(f)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_host_entry_points_match_python_binding(tmp_path):
    out = tmp_path / "o.bin"
    ply = os.path.join(ROOT, "tests", "golden", "pc_short.ply")
    _node(f"""
const c = new gt.WASDCamera({{position: [1.5, -2, 7], target: [0.25, 0, 0]}});
const raw = gt.generateScene({{n: 257, nClusters: 4, boxHalf: 3, seed: 11}});
const p = gt.loadPly({json.dumps(ply)});
let err = '';
try {{ gt.loadPly('/nonexistent/x.ply'); }} catch (e) {{ err = e.message; }}
if (!err.startsWith('gt error')) throw new Error('missing error: ' + err);
const parts = [c.matrix, raw, p.raw].map(a => Buffer.from(a.buffer, a.byteOffset, a.byteLength));
fs.writeFileSync({json.dumps(str(out))}, Buffer.concat(parts));
console.log(JSON.stringify({{n: p.n, deg: p.shDegree, version: gt.version()}}));
""", tmp_path)
    blob = np.fromfile(out, np.float32)
    m, rest = blob[:16], blob[16:]
    g, pl = rest[:257 * 59], rest[257 * 59:]
    np.testing.assert_array_equal(m, gt.camera_look_at([1.5, -2, 7], [0.25, 0, 0]))
    np.testing.assert_array_equal(g.reshape(257, 59), gt.generate_raw(n=257, n_clusters=4, box_half=3.0, seed=11))
    ref = gt.read_ply(os.path.join(ROOT, "tests", "golden", "pc_short.ply"))
    np.testing.assert_array_equal(pl.reshape(-1, 59), ref[0])


@pytest.mark.gpu
def test_render_through_addon_equals_python_binding(tmp_path):
    print("case node addon render 64x48 c1 k6", flush=True)
    out = tmp_path / "rgba.bin"
    _node(f"""
const raw = gt.generateScene({{n: 4000, boxHalf: 2, logScaleMean: -3, logScaleStd: 0.5, seed: 0}});
const tracer = new gt.GaussianTracer(raw, 4000, 0.1);
const cam = new gt.WASDCamera({{position: [0, 0, 8], target: [0, 0, 0]}});
const a = tracer.render(cam, {{alpha_min: 0.1, hit_array_size: 6, T_min: 0.03, debugMode: 0, width: 64, height: 48}});
const b = tracer.render(cam, {{alpha_min: 0.2, hit_array_size: 6, T_min: 0.03, debugMode: 0, width: 64, height: 48}});
const info = tracer.info();
const tracerGpu = new gt.GaussianTracer(raw, 4000, 0.1, {{gpuBuild: true}});
const c = tracerGpu.render(cam, {{alpha_min: 0.1, hit_array_size: 6, T_min: 0.03, debugMode: 0, width: 64, height: 48}});
tracerGpu.destroy();
tracer.destroy();
let err = '';
try {{ tracer.render(cam, {{width: 8, height: 8}}); }} catch (e) {{ err = e.message; }}
if (!err) throw new Error('render after destroy did not throw');
fs.writeFileSync({json.dumps(str(out))}, Buffer.concat([a, b, c].map(x => Buffer.from(x.buffer))));
console.log(JSON.stringify(info));
""", tmp_path)
    blob = np.fromfile(out, np.float32).reshape(3, 48, 64, 4)
    raw = gt.generate_raw(n=4000, box_half=2.0, log_scale_mean=-3.0, log_scale_std=0.5, seed=0)
    tw = gt.camera_look_at([0, 0, 8], [0, 0, 0])
    for i, amin in enumerate((0.1, 0.2)):
        s = gt.Scene(raw, alpha_min=amin)
        u = gt.make_uniforms(tw, 64, 48, hfov_deg=60, alpha_min=amin, k=6, T_min=0.03)
        want = s.render(u, 64, 48)["rgba"].reshape(48, 64, 4)
        np.testing.assert_array_equal(blob[i], want)
    assert not np.array_equal(blob[0], blob[1])
    np.testing.assert_array_equal(blob[2], blob[0])  # GPU-built scene (gpuBuild) renders the same image


@pytest.mark.gpu
def test_antialias_through_addon(tmp_path):
    print("case node addon antialiasing", flush=True)
    out = tmp_path / "aa.bin"
    _node(f"""
const raw = gt.generateScene({{n: 4000, boxHalf: 2, logScaleMean: -3, logScaleStd: 0.5, seed: 0}});
const tracer = new gt.GaussianTracer(raw, 4000, 0.1);
const cam = new gt.WASDCamera({{position: [0, 0, 8], target: [0, 0, 0]}});
const o = {{alpha_min: 0.1, hit_array_size: 6, T_min: 0.03, width: 40, height: 30, antialiasing: true}};
const f0 = tracer.render(cam, Object.assign({{}}, o, {{accumulatedFrames: 0}}));
const f1 = tracer.render(cam, Object.assign({{}}, o, {{accumulatedFrames: 1, prev: f0}}));
let err = '';
try {{ tracer.render(cam, Object.assign({{}}, o, {{accumulatedFrames: 2}})); }} catch (e) {{ err = e.message; }}
if (!err.includes('prev')) throw new Error('missing prev not reported: ' + err);
tracer.destroy();
fs.writeFileSync({json.dumps(str(out))}, Buffer.concat([f0, f1].map(x => Buffer.from(x.buffer))));
""", tmp_path)
    blob = np.fromfile(out, np.float32).reshape(2, 30, 40, 4)
    raw = gt.generate_raw(n=4000, box_half=2.0, log_scale_mean=-3.0, log_scale_std=0.5, seed=0)
    tw = gt.camera_look_at([0, 0, 8], [0, 0, 0])
    s = gt.Scene(raw, alpha_min=0.1)
    prev = None
    for frames in range(2):
        u = gt.make_uniforms(tw, 40, 30, hfov_deg=60, alpha_min=0.1, k=6, T_min=0.03, accumulated_frames=frames)
        want = s.render(u, 40, 30, antialiasing=True, rgba_prev=prev)["rgba"]
        np.testing.assert_array_equal(blob[frames], want)
        prev = want


def test_tone_map_and_png_through_addon(tmp_path):
    """toneMap / savePng (the screen pass and the Save button) equal the
    Python binding's gt_tonemap / gt_write_png bytes."""
    rng = np.random.default_rng(8)
    W, H = 37, 23
    rgba = rng.normal(0.4, 0.7, (H, W, 4)).astype(np.float32)
    src = tmp_path / "in.bin"
    rgba.tofile(src)
    _node(f"""
const rgba = new Float32Array(fs.readFileSync({json.dumps(str(src))}).buffer.slice(0));
for (const mode of [0, 1, 2]) {{
  fs.writeFileSync({json.dumps(str(tmp_path))} + '/m' + mode + '.bin', Buffer.from(gt.toneMap(rgba, {W}, {H}, mode)));
}}
gt.savePng({json.dumps(str(tmp_path / 'js.png'))}, rgba, {W}, {H}, 1);
let err = '';
try {{ gt.toneMap(rgba, {W} + 1, {H}); }} catch (e) {{ err = e.message; }}
if (!err.includes('width * height * 4')) throw new Error('missing size error: ' + err);
""", tmp_path)
    for mode in (0, 1, 2):
        js = np.fromfile(tmp_path / f"m{mode}.bin", np.uint8).reshape(H, W, 4)
        np.testing.assert_array_equal(js, gt.tonemap(rgba, mode))
    gt.write_png(tmp_path / "py.png", gt.tonemap(rgba, 1))
    assert (tmp_path / "js.png").read_bytes() == (tmp_path / "py.png").read_bytes()


@pytest.mark.gpu
def test_render_into_reused_array(tmp_path):
    """opts.out: the frame is written into (and returned as) the caller's
    Float32Array; it equals a render into a new array bit for bit."""
    _node("""
const raw = gt.generateScene({n: 3000, nClusters: 2, boxHalf: 2, seed: 5});
const tr = new gt.GaussianTracer(raw, 3000, 0.1);
const cam = new gt.WASDCamera({position: [0, 0, 8], target: [0, 0, 0]});
const opts = {width: 96, height: 64, alpha_min: 0.1, hit_array_size: 6, T_min: 0.03};
const a = tr.render(cam, opts);
const buf = new Float32Array(96 * 64 * 4).fill(7);
for (let i = 0; i < 2; ++i) {
  const b = tr.render(cam, Object.assign({out: buf}, opts));
  if (b !== buf) throw new Error('out not returned');
}
for (let i = 0; i < a.length; ++i) if (a[i] !== buf[i]) throw new Error('differs at ' + i);
let err = '';
try { tr.render(cam, Object.assign({out: new Float32Array(5)}, opts)); } catch (e) { err = e.message; }
if (!err.includes('opts.out')) throw new Error('missing size error: ' + err);
tr.destroy();
""", tmp_path)
