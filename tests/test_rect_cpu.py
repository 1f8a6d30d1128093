"""CPU test of the proxy screen rectangles that the packet traversal culls
leaf proxies with (webgpu-gaussiantracer_amd/csrc/gt_rect.h).

tests/rect/check_rect.c, built here with gcc (FP32 without contraction, as
the kernels are built), draws random proxies (anisotropic scales over five
decades, distances 0.05-2000, centres ahead of, around and behind the
camera) and pinhole frames (fields of view 1-120 degrees, scenes offset
1e4 from the world origin, non-orthogonal bases, jittered and centred pixel
offsets, t_min 0 and 1e-3).  For every pixel it evaluates the FP32
bounding-sphere stage of the proxy test as the kernel does, and requires
each accepted pixel to lie in the proxy's rectangle; for proxies whose
rectangle is empty (wholly behind the camera) it runs the exact triangle
test instead and requires no hit.  The same rectangles shrunk by 3 pixels
must miss accepted pixels (the check sees the boundary), and the 15-bit
field overlap test must agree with plain comparisons.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(tmp_path, cams, prox):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path / "check_rect")
    subprocess.run(["gcc", "-std=gnu99", "-O2", "-ffp-contract=off", "-Wall", "-Werror", "-Wno-unused-function",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "webgpu-gaussiantracer_amd", "csrc"),
                    os.path.join(ROOT, "tests", "rect", "check_rect.c"), "-lm", "-o", exe], check=True)
    out = subprocess.run([exe, str(cams), str(prox)], capture_output=True, text=True, timeout=300)
    vals = dict((k, int(v)) for k, v in re.findall(r"(\w+) (\d+)", out.stdout))
    return out, vals


def test_rectangles_hold_every_accepted_ray(tmp_path):
    out, v = _run(tmp_path, 40, 400)
    assert out.returncode == 0, out.stdout + out.stderr
    assert v["misses"] == 0 and v["swar_bad"] == 0
    assert v["accepted"] > 10_000_000 and v["behind"] > 50_000
    assert v["full"] > 0 and v["empty"] > 0
    assert v["shrunk"] > 0
    # the rectangles cull: they cover well under the frame area summed over proxies
    assert v["inside"] < 0.75 * v["cases"]
