"""Pins the oracle's DenseUpsampleLayer (DenseUpsampleLayer.cs) and its rcpss model (IntelApprox.cs:20-28)."""
import numpy as np
import pytest

from oracle import oracle as o


def _intel():
    try:
        return any(l.startswith("vendor_id") and "GenuineIntel" in l for l in open("/proc/cpuinfo"))
    except OSError:
        return False


@pytest.mark.skipif(not _intel(), reason="the rcpss table is Intel's: checked against the instruction on an Intel host")
def test_rcp_model_matches_the_instruction():
    rng = np.random.default_rng(0)
    bits = list(rng.integers(0, 2 ** 32, 200000, dtype=np.uint64))
    special = [0x0, 0x80000000, 0x7f800000, 0xff800000, 0x7fc00000, 0x7f800001, 0x00000001, 0x807fffff, 0x00800000,
               0x7f7fffff, 0xff7fffff, 0x3f800000, 0x7e800000, 0x7e7fffff, 0x3fffffff]
    for e in range(256):  # every exponent, a few mantissas
        for m in (0, 1, 0x7ff, 0x800, 0x400000, 0x7fffff):
            special.append((e << 23) | m)
            special.append(0x80000000 | (e << 23) | m)
    for b in bits + special:
        assert o.rcp_bits(b) == o.rcpss_hw(b), hex(int(b))


def test_rcp_special_cases():
    assert o.rcp_bits(0x00000000) == 0x7f800000 and o.rcp_bits(0x80000000) == 0xff800000  # +-0 -> +-inf
    assert o.rcp_bits(0x00000005) == 0x7f800000                                           # denormal -> inf
    assert o.rcp_bits(0x7f800000) == 0 and o.rcp_bits(0xff800000) == 0x80000000            # +-inf -> +-0
    assert o.rcp_bits(0x7f800001) == 0x7fc00001                                           # NaN -> quiet NaN
    assert o.rcp_bits(0x7f7fffff) == 0                                                    # result would be denormal
    v = np.array([o.rcp_bits(np.float32(x).view(np.uint32)) for x in (1.0, 2.0, 0.5, 3.0)], np.uint32).view(np.float32)
    np.testing.assert_allclose(v, [1.0, 0.5, 2.0, 1 / 3], rtol=4e-4)  # |rel err| <= 1.5 * 2^-12


def test_zero_shift_is_identity_and_half_shift_is_cubic():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (9, 13, 4)).astype(np.uint8)
    np.testing.assert_array_equal(o.image_shift(img, 0.0, 0.0), img)
    flat = np.full((5, 6, 4), 77, np.uint8)
    np.testing.assert_array_equal(o.image_shift(flat, 0.5, 0.5), flat)  # the taps sum to 1: a flat image stays flat
    ramp = np.repeat(np.repeat(np.arange(16, dtype=np.uint8)[None, :, None] * 10, 3, 0), 4, 2)
    out = o.image_shift(ramp, 0.5, 0.0)
    np.testing.assert_array_equal(out[:, 2:-3, 0], ramp[:, 2:-3, 0] + 5)  # a cubic is exact on a line: x + 0.5


def test_nearest_resize_bruteforce():
    rng = np.random.default_rng(2)
    for (sh, sw, H, W) in ((8, 12, 4, 6), (7, 9, 3, 5), (3, 3, 5, 7)):
        img = rng.integers(0, 256, (sh, sw, 4)).astype(np.uint8)
        out = o.nearest_resize(img, W, H)
        sy_step, sx_step = int(sh * 65536.0 / H), int(sw * 65536.0 / W)
        for y in range(H):
            for x in range(W):
                np.testing.assert_array_equal(out[y, x], img[(y * sy_step) >> 16, (x * sx_step) >> 16])


def _f(b):
    return np.array([b], np.uint32).view(np.float32)[0]


def _exp2_twin(t):
    f32 = np.float32
    t = f32(t) if t < 128 else f32(128)
    t = t if t > -126 else f32(-126)
    n = int(t) + (-1 if np.signbit(t) else 0)
    fr = f32(t - f32(n))
    p = f32(f32(f32(f32(f32(fr * _f(0x3d9fcb52)) + _f(0x3e677e26)) * fr) + _f(0x3f322226)) * fr) + _f(0x3f7ffb19)
    return np.array([(int(np.float32(p).view(np.uint32)) + (n << 23)) & 0xFFFFFFFF], np.uint32).view(np.float32)[0]


def _bilateral_twin(lo, guide, g2, sigma):
    """float32 twin of BilateralUpsampleFromCollapse, one rounded operation at a time (tiny images only)."""
    f32 = np.float32
    H, W = guide.shape[:2]
    lh, lw = lo.shape
    k = f32(f32(0.5) / f32(f32(sigma) * f32(sigma)))
    eps, third = _f(0x322bcc77), _f(0x3eaaaaab)

    def sw(fv, i):
        d = abs(f32(fv - f32(i * 2)))
        d = d if d < 1 else f32(1)
        return f32(1) if int(d) == 0 else third

    out = np.zeros((H, W), np.float32)
    for y in range(H):
        iy0 = int((y - 1) / 2)
        fy = f32(f32(y) + f32(-0.5))
        for x in range(W):
            hr, hg, hb = (f32(v) for v in guide[y, x, :3])
            ix0 = int((x - 1) / 2)
            fx = f32(f32(x) + f32(-0.5))
            ibx, ix1 = min(max(ix0, 0), lw - 1), min(max(ix0 + 1, 0), lw - 1)
            wx0, wx1 = sw(fx, ibx), sw(fx, ix1)
            av, as_ = f32(0), f32(0)
            for j in range(2):
                iy = min(max(iy0 + j, 0), lh - 1)
                wy = sw(fy, iy)
                ws = []
                for ix, wx in ((ibx, wx0), (ix1, wx1)):
                    d0, d1, d2 = (f32(h - f32(g)) for h, g in zip((hr, hg, hb), g2[iy, ix, :3]))
                    s = f32(f32(d1 * d1) + f32(f32(d2 * d2) + f32(d0 * d0)))
                    ws.append(f32(f32(_exp2_twin(-f32(s * k)) + eps) * f32(wy * wx)))
                v0, v1 = f32(lo[iy, ibx] * ws[0]), f32(lo[iy, ix1] * ws[1])
                av = f32(v1 + f32(v0 + av))
                as_ = f32(ws[1] + f32(ws[0] + as_))
            out[y, x] = f32(av / as_)
    return out


def test_bilateral_matches_float32_twin():
    rng = np.random.default_rng(3)
    lh, lw = 4, 5
    lo = rng.uniform(0.001, 0.01, (lh, lw)).astype(np.float32)
    guide = rng.integers(0, 256, (2 * lh, 2 * lw, 4)).astype(np.uint8)
    g2 = rng.integers(0, 256, (lh, lw, 4)).astype(np.uint8)
    g2[0, 0, :3] = guide[0, 0, :3]  # a zero range difference (t = -0: the sign-bit exponent step)
    np.testing.assert_array_equal(o.bilateral_upsample(lo, guide, g2, 12.0), _bilateral_twin(lo, guide, g2, 12.0))


def test_upsample_layer_size_rule_and_flat_depth():
    rng = np.random.default_rng(4)
    guide = rng.integers(0, 256, (12, 16, 4)).astype(np.uint8)
    depth = np.full((6, 8), 1500.0, np.float32)
    out = o.upsample_layer(depth, guide, 8, 6)
    assert out.shape == (12, 16)
    np.testing.assert_allclose(out, 1500.0, rtol=1e-3)  # rcp twice (~2^-12 each) of a flat depth
    with pytest.raises(RuntimeError):
        o.upsample_layer(depth, guide, 8, 5)  # "Incorrect output depth size!"


def test_bilateral_follows_the_one_matching_tap():
    """A structural pin that does not restate the arithmetic: every low-res pixel has its own colour, 255 apart from
    the others' on some channel, so for an output pixel whose guide colour equals exactly one of its four taps the
    range weights of the other three are 2^-126 + 1e-8 against 1, and the result is that tap's value to float32
    precision. Which tap it is follows (x-1)/2, (x-1)/2 + 1 (clamped) and the same in y; a wrong index, a swapped
    axis or a dropped spatial weight moves it to another tap."""
    lw, lh = 6, 5
    rng = np.random.default_rng(3)
    lo = rng.uniform(1.0, 9.0, (lh, lw)).astype(np.float32)
    colours = np.zeros((lh, lw, 4), np.uint8)
    for j in range(lh):
        for i in range(lw):  # corners of the RGB cube, distinct for neighbouring pixels (parities of i and j)
            colours[j, i] = (255 * (i % 2), 255 * (j % 2), 255 * ((i // 2 + j // 2) % 2), 0)
    W, H = 2 * lw, 2 * lh
    guide = np.zeros((H, W, 4), np.uint8)
    want = np.zeros((H, W), np.float32)
    for y in range(H):
        iy0 = (y - 1) // 2 if y >= 1 else 0   # C division toward zero: (y-1)/2 is 0 for y = 0
        for x in range(W):
            ix0 = (x - 1) // 2 if x >= 1 else 0
            taps = [(min(max(iy0 + dj, 0), lh - 1), min(max(ix0 + di, 0), lw - 1)) for dj in (0, 1) for di in (0, 1)]
            ty, tx = taps[(x + 2 * y) % 4]    # pick one of the four taps per pixel
            guide[y, x] = colours[ty, tx]
            want[y, x] = lo[ty, tx]
    out = o.bilateral_upsample(lo, guide, colours, 12.0)
    # every chosen tap is the only one of the four with that colour, except where clamping repeats a tap
    np.testing.assert_allclose(out, want, rtol=1e-6)
