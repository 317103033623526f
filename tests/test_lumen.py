"""Lumen's dump format (StereoAsyncApi.cs:196, :212-248) and the depthOverride seam: byte layout on a hand-built
layer, round trips, and the depth closed form (no device needed)."""
import numpy as np

from lux_amd import lumen


def test_costvol_byte_layout(tmp_path):
    prefix = str(tmp_path / "L0")
    # pixel 0 computed: start 3, count 5, cap 8; pixel 1 skipped: start 0, count 2, cap 8
    agg = np.arange(16, dtype=np.uint16) + 1000
    raw = np.arange(16, dtype=np.uint8)
    lumen.write_costvol(prefix, 2, 1, [3, 0], [5, 2], 8, agg, raw=raw, skip=[0, 0xFF])
    ob = open(prefix + "_offsets.u32", "rb").read()
    assert ob == np.array([2, 1, 2, 4], "<i4").tobytes() + np.array([0, 8 + 16 + 8], "<u4").tobytes()
    cb = open(prefix + "_costvol.bin", "rb").read()
    want = (np.array([3, 5, 1, 8], "<u2").tobytes() + agg[:8].astype("<u2").tobytes() + raw[:8].tobytes()
            + np.array([0, 2, 1, 8], "<u2").tobytes() + agg[8:].astype("<u2").tobytes())
    assert cb == want
    d = lumen.read_costvol(prefix)
    assert d["W"] == 2 and d["H"] == 1
    np.testing.assert_array_equal(d["start"], [3, 0])
    np.testing.assert_array_equal(d["count"], [5, 2])
    np.testing.assert_array_equal(d["skip"], [0, 0xFF])
    np.testing.assert_array_equal(d["agg"], agg)
    np.testing.assert_array_equal(d["raw"][:8], raw[:8])
    np.testing.assert_array_equal(d["raw"][8:], [0, 0, 255, 255, 255, 255, 255, 255])


def test_costvol_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    W, H, align = 7, 5, 2
    start = rng.integers(0, 30, W * H).astype(np.uint16)
    count = rng.integers(0, 9, W * H).astype(np.uint16)
    cap = lumen.caps_of(count, align)
    agg = rng.integers(0, 65536, int(cap.sum())).astype(np.uint16)
    agg1 = rng.integers(0, 65536, int(cap.sum())).astype(np.uint16)
    raw = rng.integers(0, 256, int(cap.sum())).astype(np.uint8)
    prefix = str(tmp_path / "L3")
    lumen.write_costvol(prefix, W, H, start, count, align, agg, raw=raw, agg_pass1=agg1)
    d = lumen.read_costvol(prefix)
    np.testing.assert_array_equal(d["agg"], agg)
    np.testing.assert_array_equal(d["raw"], raw)
    np.testing.assert_array_equal(lumen.read_costvol(prefix, pass1=True)["agg"], agg1)


def test_images_and_depth_seam(tmp_path):
    rgba = np.random.default_rng(1).integers(0, 256, (3, 4, 4)).astype(np.uint8)
    lumen.write_image(str(tmp_path / "g.img"), rgba)
    b = open(tmp_path / "g.img", "rb").read()
    assert b[:16] == np.array([4, 3, 4, 4], "<i4").tobytes()
    np.testing.assert_array_equal(lumen.read_image(str(tmp_path / "g.img"), np.uint8), rgba)
    sub = np.array([[2.0, 0.5], [0.0, 4.0]], np.float32)
    valid = np.array([[1, 1], [1, 0]], np.uint8)
    depth = lumen.depth_override(sub, valid, 100.0, invalid=-1.0)
    np.testing.assert_array_equal(depth, np.array([50.0, 200.0, -1.0, -1.0], np.float32))
    lumen.write_depth(str(tmp_path / "d.f32"), depth, 2, 2)
    np.testing.assert_array_equal(lumen.read_image(str(tmp_path / "d.f32"), np.float32).reshape(-1), depth)
