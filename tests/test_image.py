"""Presentation of a frame (gt_tonemap, gt_write_png): the reference's screen
pass (src/toneMapper.ts:11-26) and Save button (src/main.ts:606-611).

Host-only entry points: no GPU.  The tone map is pinned by values the
definitions fix (unorm steps, the exposure curve at ln 2 / 1.66, gamma of
0.5); the PNG files are read back by an independent decoder (Python's zlib,
which checks the Adler-32, and zlib.crc32 for every chunk)."""
import math
import struct
import zlib

import numpy as np
import pytest

import gtrace as gt


def _frame(values):
    v = np.asarray(values, np.float32)
    rgba = np.zeros((1, v.size, 4), np.float32)
    rgba[0, :, 0] = v
    rgba[0, :, 1] = v
    rgba[0, :, 2] = v
    rgba[0, :, 3] = 0.25  # ignored: the canvas alpha is 1
    return rgba


def test_identity_unorm_steps():
    vals = [0.0, 1.0, 0.5, 1.0 / 255.0, 0.5 / 255.0 * 0.999, 0.5 / 255.0 * 1.001, 254.5 / 255.0 * 1.0001,
            -1.0, 2.0, math.inf, -math.inf, math.nan, 1e-40]
    want = [0, 255, 128, 1, 0, 1, 255, 0, 255, 255, 0, 0, 0]
    out = gt.tonemap(_frame(vals))
    assert out[0, :, 0].tolist() == want
    assert (out[0, :, 1] == out[0, :, 0]).all() and (out[0, :, 2] == out[0, :, 0]).all()
    assert (out[..., 3] == 255).all()


def test_exposure_and_gamma_fixed_points():
    k = float(np.float32(1.66))
    x = -math.log(0.75) / k  # 1 - exp(-k x) = 1/4: 63.75 steps, away from a rounding tie
    out1 = gt.tonemap(_frame([0.0, x, 50.0, -0.1]), gt.TONEMAP_EXPOSURE)[0, :, 0].tolist()
    assert out1 == [0, 64, 255, 0]
    out2 = gt.tonemap(_frame([0.0, x, 50.0]), gt.TONEMAP_EXPOSURE_GAMMA)[0, :, 0].tolist()
    assert out2 == [0, 136, 255]  # 255 * 0.25 ** (1 / 2.2) = 135.79


def test_random_frame_against_numpy():
    rng = np.random.default_rng(4)
    rgba = rng.normal(0.4, 0.6, (61, 77, 4)).astype(np.float32)
    x = rgba[..., :3].astype(np.float64)
    k = float(np.float32(1.66))
    curves = {gt.TONEMAP_IDENTITY: x, gt.TONEMAP_EXPOSURE: 1.0 - np.exp(-k * x)}
    curves[gt.TONEMAP_EXPOSURE_GAMMA] = np.where(curves[1] > 0, np.maximum(curves[1], 1e-300) ** (1 / 2.2), 0.0)
    for mode, c in curves.items():
        want = np.where(c > 0, np.where(c >= 1, 255, np.floor(c * 255.0 + 0.5)), 0).astype(np.int64)
        got = gt.tonemap(rgba, mode)[..., :3].astype(np.int64)
        diff = np.abs(got - want)
        assert diff.max() <= 1, mode
        assert (diff == 0).mean() > 0.9999, mode  # only ties of an exp/log ulp may round apart


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body), typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    assert len(raw) == (w * 4 + 1) * h
    rows = np.frombuffer(raw, np.uint8).reshape(h, w * 4 + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (300, 250), (16383, 1), (16384, 1)])
def test_png_round_trip(tmp_path, w, h):
    img = np.random.default_rng(w * 31 + h).integers(0, 256, (h, w, 4), dtype=np.uint8)
    p = tmp_path / "a.png"
    gt.write_png(p, img)
    np.testing.assert_array_equal(_read_png(p), img)


def test_rendered_frame_saved_as_the_canvas_shows_it(tmp_path):
    """Row y of the frame is canvas row y (pixelIndex = x + y W, toneMapper.ts:16)."""
    rgba = np.zeros((2, 3, 4), np.float32)
    rgba[0, 0, :3] = (1.0, 0.0, 0.0)
    rgba[1, 2, :3] = (0.0, 0.0, 1.0)
    p = tmp_path / "f.png"
    gt.write_png(p, gt.tonemap(rgba))
    back = _read_png(p)
    assert back[0, 0].tolist() == [255, 0, 0, 255]
    assert back[1, 2].tolist() == [0, 0, 255, 255]
    assert back[0, 1].tolist() == [0, 0, 0, 255]


def test_errors(tmp_path):
    with pytest.raises(gt.GTError):
        gt.tonemap(_frame([0.5]), mode=3)
    with pytest.raises(gt.GTError):
        gt.write_png(tmp_path / "no" / "dir" / "x.png", np.zeros((1, 1, 4), np.uint8))
    with pytest.raises(gt.GTError):
        gt.write_png(tmp_path / "big.png", np.zeros((1, 16385, 4), np.uint8))
    empty = gt.tonemap(np.zeros((0, 5, 4), np.float32))
    assert empty.shape == (0, 5, 4)
