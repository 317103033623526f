"""Pins of the oracle's plane-sweep matching cost (DenseStereo.cs:217-381) and the layer set-up around it
(DenseStereoSgm.cs:34-140, :225-231): closed forms chosen so that a dropped term, a wrong index, a swapped operand
or a wrong rounding fails one of them, plus a numpy twin vectorised over pixels for random warps."""
import numpy as np
import pytest

from oracle import oracle as o

F27 = float(np.array([0x3d17b426], np.uint32).view(np.float32)[0])
COLOUR = (8160, 680, 680, 0)
GRAY = (12240, 0, 0, 0)


def translation_warp(tx, ty):
    """A flow matrix whose projection is (x + tx, y + ty) at every depth: p = z*(x + tx, y + ty, 1)."""
    m = np.zeros(16, np.float32)
    m[0] = m[5] = m[10] = 1.0      # column 0 = (1,0,0,0), column 1 = (0,1,0,0), column 2 = (tx, ty, 1, 0)
    m[8], m[9] = tx, ty
    m[15] = 1.0
    return m


def baseline_warp(fb):
    """Rectified horizontal pair: x' = x - fb / z."""
    m = np.zeros(16, np.float32)
    m[0] = m[5] = m[10] = m[15] = 1.0
    m[12] = -fb
    return m


def planes(n, near=400.0, far=4000.0):
    inv = np.linspace(1.0 / far, 1.0 / near, n, dtype=np.float64)
    return (1.0 / inv).astype(np.float32)


def textured(h, w, seed):
    rng = np.random.default_rng(seed)
    base = rng.integers(20, 230, (h // 2 + 2, w // 2 + 2, 4)).astype(np.uint8)
    return np.repeat(np.repeat(base, 2, 0), 2, 1)[:h, :w].copy()


def shifted(img, dx, dy):
    """out(x, y) = img(x - dx, y - dy) where defined, else a constant."""
    h, w = img.shape[:2]
    out = np.full_like(img, 77)
    ys, xs = slice(max(dy, 0), h + min(dy, 0)), slice(max(dx, 0), w + min(dx, 0))
    out[ys, xs] = img[max(-dy, 0):h - max(dy, 0), max(-dx, 0):w - max(dx, 0)]
    return out


def test_matching_cost_is_zero_on_the_true_shift():
    ref = textured(40, 48, 1)
    other = shifted(ref, 3, -2)                   # other(x + 3, y - 2) = ref(x, y)
    sc = o.SweepContext(ref, [other], [translation_warp(3.0, -2.0)], [COLOUR], planes(16))
    for x, y in ((10, 10), (20, 30), (5, 4)):
        np.testing.assert_array_equal(o.matching_cost(sc, x, y, 0, 16), 0)


def test_matching_cost_constant_offset_closed_form():
    """other = ref + 1 on every channel: every capped difference is 1 (caps >= 1) or 0 (cap 0), so per camera
    ((9*8160+16)>>5) + 2*((9*680+16)>>5) = 2295 + 2*191 = 2677 (colour) and (9*12240+16)>>5 = 3443 (gray)."""
    ref = textured(30, 30, 2)
    other = ref + 1
    sc = o.SweepContext(ref, [other], [translation_warp(0.0, 0.0)], [COLOUR], planes(8))
    np.testing.assert_array_equal(o.matching_cost(sc, 12, 12, 0, 8), 2677)
    sc2 = o.SweepContext(ref, [other, other], [translation_warp(0.0, 0.0)] * 2, [COLOUR, GRAY], planes(8))
    np.testing.assert_array_equal(o.matching_cost(sc2, 12, 12, 2, 5), 2677 + 3443)


def test_matching_cost_caps_and_channel_order():
    """A difference of 10 on one channel at a time: capped to caps[ch], weighted by Wq[ch]."""
    ref = textured(30, 30, 3)
    for ch, (cap, wq) in enumerate(zip((2, 6, 6, 0), COLOUR)):
        other = ref.astype(np.int32)
        other[:, :, ch] = np.clip(other[:, :, ch] + 10, 0, 255)
        sc = o.SweepContext(ref, [other.astype(np.uint8)], [translation_warp(0.0, 0.0)], [COLOUR], planes(4))
        want = (9 * cap * wq + 16) >> 5
        np.testing.assert_array_equal(o.matching_cost(sc, 15, 15, 0, 4), want, err_msg=f"channel {ch}")


def test_matching_cost_half_pixel_averages():
    """A horizontal ramp 2x + c: pavgb of neighbours is 2x + c + 1, the ramp half a pixel on. A warp of +0.25 puts
    fx = x + 0.5, so hx = 1 and the averaged block matches a reference that is the ramp plus one. The planes are
    powers of two so that the projection is exact (x + 0.5 sits on the rounding edge of hx)."""
    pow2 = np.float32(2.0) ** np.arange(4, 10, dtype=np.float32)
    h, w = 24, 60
    x = np.arange(w)
    ramp = np.zeros((h, w, 4), np.uint8)
    for ch, c in enumerate((10, 30, 50, 70)):
        ramp[:, :, ch] = 2 * x + c
    ref = ramp + 1
    sc = o.SweepContext(ref, [ramp], [translation_warp(0.25, 0.0)], [COLOUR], pow2)
    np.testing.assert_array_equal(o.matching_cost(sc, 20, 10, 0, 6), 0)
    # vertical ramp, half a pixel in y
    y = np.arange(h)
    vr = np.zeros((h, w, 4), np.uint8)
    for ch, c in enumerate((10, 30, 50, 70)):
        vr[:, :, ch] = (2 * y + c)[:, None]
    sc = o.SweepContext(vr + 1, [vr], [translation_warp(0.0, 0.25)], [COLOUR], pow2)
    np.testing.assert_array_equal(o.matching_cost(sc, 20, 10, 0, 6), 0)
    # without the half pixel the ramp + 1 reference differs by 1 everywhere: 2677
    sc = o.SweepContext(ref, [ramp], [translation_warp(0.0, 0.0)], [COLOUR], pow2)
    np.testing.assert_array_equal(o.matching_cost(sc, 20, 10, 0, 6), 2677)


def test_matching_cost_wraps_over_cameras_and_clamps_per_camera():
    ref = np.full((20, 20, 4), 100, np.uint8)
    other = np.full((20, 20, 4), 200, np.uint8)
    caps = (28, 0, 0, 0)
    one = (9 * 28 * 5080 + 16) >> 5          # 40005 per camera
    sc = o.SweepContext(ref, [other, other], [translation_warp(0.0, 0.0)] * 2, [(5080, 0, 0, 0)] * 2, planes(3), caps)
    np.testing.assert_array_equal(o.matching_cost(sc, 8, 8, 0, 3), (2 * one) % 65536)
    sc = o.SweepContext(ref, [other], [translation_warp(0.0, 0.0)], [(12240, 0, 0, 0)], planes(3), caps)
    np.testing.assert_array_equal(o.matching_cost(sc, 8, 8, 0, 3), 65535)   # 96390 clamps


def test_matching_cost_carries_the_last_cost_out_of_the_rect():
    """x' = x - fb/z leaves the rect [1, w-3] for near planes: those planes repeat the last in-rect cost; an
    in-rect plane equals its own single-plane evaluation; the first plane is clamped into the rect."""
    ref = textured(32, 64, 4)
    other = textured(32, 64, 5)
    pl = planes(40, near=300.0, far=3000.0)
    sc = o.SweepContext(ref, [other], [baseline_warp(9000.0)], [COLOUR], pl)
    x, y = 15, 12
    full = o.matching_cost(sc, x, y, 0, 40)
    fx = x - 9000.0 / pl.astype(np.float64) + 0.25
    inside = (np.floor(fx) >= 1) & (np.floor(fx) < 64 - 2)
    assert inside[0] and not inside.all()
    last = None
    for d in range(40):
        if inside[d]:
            assert full[d] == o.matching_cost(sc, x, y, d, 1)[0]
            last = full[d]
        else:
            assert full[d] == last, d
    # a range starting out of the rect: its first plane is clamped to the rect's left column
    first_out = int(np.argmin(inside))
    edge = o.SweepContext(ref, [other], [translation_warp(1.0 - x, 0.0)], [COLOUR], pl)   # fx = 1.25 exactly
    assert o.matching_cost(sc, x, y, first_out, 1)[0] == o.matching_cost(edge, x, y, 0, 1)[0]


def test_normalise_closed_forms():
    assert F27 == np.float32(1.0 / 27.0)
    raw = o.normalise(np.array([27, 0, 26, 65535, 270], np.uint16), 5, 8, 1)
    # 27/27 -> 1; 26/27 -> 0 (truncation); 65535/27 = 2427.2 -> 2427 & 255 = 123; padding 255
    np.testing.assert_array_equal(raw, [1, 0, 0, 123, 10, 255, 255, 255])
    raw = o.normalise(np.array([54 * 3, 53 * 3], np.uint16), 2, 2, 3)
    np.testing.assert_array_equal(raw, [2, 1])


def _numpy_layer_cost(ref, others, warps, wqs, pl, caps, W, H, scale, start, count, align):
    """The layer's raw volume written from the reference's description with numpy, vectorised over pixels: float32
    arithmetic one rounded operation at a time, saturating conversions, the carried cost as a running select."""
    rh, rw = ref.shape[:2]
    half = scale // 2
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    xi = np.minimum(scale * xs + half, rw - 1).reshape(-1)
    yi = np.minimum(scale * ys + half, rh - 1).reshape(-1)
    npx = W * H
    start = start.reshape(-1).astype(np.int64)
    count = count.reshape(-1).astype(np.int64)
    cmax = int(count.max())
    total = np.zeros((npx, cmax), np.int64)
    refb = np.zeros((npx, 3, 3, 4), np.int32)
    for r in range(3):
        for c in range(3):
            refb[:, r, c] = ref[np.clip(yi + r - 1, 0, rh - 1), np.clip(xi + c - 1, 0, rw - 1)]

    def to_int(v):
        v = np.asarray(v, np.float32)
        out = np.where(np.isnan(v), 0, np.clip(np.nan_to_num(v, nan=0.0), -2.0 ** 31, 2.0 ** 31 - 1))
        return np.trunc(out).astype(np.int64)

    f32 = np.float32
    for img, m, wq in zip(others, warps, wqs):
        m = np.asarray(m, np.float32)
        ih, iw = img.shape[:2]
        x0, y0, x1, y1 = 1, 1, iw - 2, ih - 2
        last = np.zeros(npx, np.int64)
        lpos = np.full((npx, 4), -(1 << 40), np.int64)
        imgi = img.astype(np.int32)
        for d in range(cmax):
            act = d < count
            z = pl[np.minimum(start + d, len(pl) - 1)].astype(np.float32)
            xz, yz = (xi.astype(f32) * z).astype(f32), (yi.astype(f32) * z).astype(f32)
            def lin(a, b, c0, c1):
                return (((z * m[a]).astype(f32) + m[b]).astype(f32) + (xz * m[c0]).astype(f32)).astype(f32) + (yz * m[c1]).astype(f32)
            px, py, pz = lin(8, 12, 0, 4).astype(f32), lin(9, 13, 1, 5).astype(f32), lin(10, 14, 2, 6).astype(f32)
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = (f32(1.0) / pz).astype(f32)
                fx = ((px * inv).astype(f32) + f32(0.25)).astype(f32)
                fy = ((py * inv).astype(f32) + f32(0.25)).astype(f32)
            if d == 0:
                fx = np.fmin(np.fmax(fx, f32(x0)), f32(x1 - 1))
                fy = np.fmin(np.fmax(fy, f32(y0)), f32(y1 - 1))
            ix, iy = to_int(fx), to_int(fy)
            hx, hy = to_int((fx + fx).astype(f32)) & 1, to_int((fy + fy).astype(f32)) & 1
            pos = np.stack([ix, iy, hx, hy], 1)
            inside = (ix >= x0) & (ix < x1) & (iy >= y0) & (iy < y1)
            same = (pos == lpos).all(1)
            fresh = inside & ~same & act
            c = np.zeros(npx, np.int64)
            if fresh.any():
                k = np.nonzero(fresh)[0]
                S = np.zeros((len(k), 4), np.int64)
                for q in range(3):
                    for p in range(3):
                        a = imgi[iy[k] - 1 + q, ix[k] - 1 + p]
                        b = imgi[iy[k] - 1 + q, ix[k] + p]
                        hyk = hy[k][:, None] == 1
                        a = np.where(hyk, (a + imgi[iy[k] + q, ix[k] - 1 + p] + 1) >> 1, a)
                        b = np.where(hyk, (b + imgi[iy[k] + q, ix[k] + p] + 1) >> 1, b)
                        t = np.where(hx[k][:, None] == 1, (a + b + 1) >> 1, a)
                        S += np.minimum(np.abs(t - refb[k, q, p]), np.array(caps))
                cc = sum((S[:, ch] * wq[ch] + 16) >> 5 for ch in range(4))
                c[k] = np.minimum(cc, 65535)
            c = np.where(fresh, c, last)
            total[:, d] += np.where(act, c, 0)
            last = np.where(fresh, c, last)
            lpos = np.where(fresh[:, None], pos, lpos)
    cost = (total & 0xFFFF).astype(np.float32)
    f = np.float32(F27) / np.float32(len(others))
    raw8 = (to_int((cost * f).astype(np.float32)) & 0xFF).astype(np.uint8)
    out = []
    for i in range(npx):
        cap = -(-int(count[i]) // align) * align
        r = np.full(cap, 255, np.uint8)
        r[:count[i]] = raw8[i, :count[i]]
        out.append(r)
    return np.concatenate(out)


@pytest.mark.parametrize("scale,nv,ranged", [(4, 1, False), (2, 3, True), (8, 2, True)])
def test_layer_sweep_cost_matches_numpy_twin(scale, nv, ranged):
    rng = np.random.default_rng(scale * 10 + nv)
    H0, W0 = 48, 64
    ref = textured(H0, W0, 10)
    others, warps = [], []
    for k in range(nv):
        others.append(textured(H0 + 4 * k, W0 - 2 * k, 20 + k))
        m = baseline_warp(float(rng.uniform(2000, 9000)))
        m[1], m[4] = rng.uniform(-0.02, 0.02, 2)                  # a small rotation-like shear
        m[2], m[6] = rng.uniform(-1e-4, 1e-4, 2)                  # and a perspective term
        m[13] = rng.uniform(-300, 300)
        warps.append(m)
    wqs = [COLOUR if k % 2 == 0 else GRAY for k in range(nv)]
    pl = planes(24, near=500.0, far=5000.0)
    W, H = -(-W0 // scale), -(-H0 // scale)
    if ranged:
        start = rng.integers(0, 18, (H, W)).astype(np.uint16)
        count = rng.integers(0, 24 - start + 1).astype(np.uint16)
        count[0, 0] = 0
    else:
        start = np.zeros((H, W), np.uint16)
        count = np.full((H, W), 24, np.uint16)
    sc = o.SweepContext(ref, others, warps, wqs, pl)
    got = o.layer_sweep_cost(sc, W, H, scale, 8, start, count)
    want = _numpy_layer_cost(ref, others, warps, wqs, pl, (2, 6, 6, 0), W, H, scale, start, count, 8)
    np.testing.assert_array_equal(got, want)


def test_layer_sweep_cost_skipped_pixels():
    ref = textured(16, 16, 7)
    sc = o.SweepContext(ref, [ref], [translation_warp(0.0, 0.0)], [COLOUR], planes(5))
    skip = np.zeros((4, 4), np.uint8)
    skip[1, 2] = 0xFF
    got = o.layer_sweep_cost(sc, 4, 4, 4, 8, np.zeros((4, 4)), np.full((4, 4), 5), skip).reshape(16, 8)
    np.testing.assert_array_equal(got[6], [0, 0, 0, 0, 0, 255, 255, 255])
    np.testing.assert_array_equal(got[5], [0, 0, 0, 0, 0, 255, 255, 255])   # the true match costs 0 too


def test_guidance_subsample():
    ref = np.random.default_rng(8).integers(0, 256, (50, 70, 4)).astype(np.uint8)
    g = o.guidance(ref, 9, 7, 8)
    for y in range(7):
        for x in range(9):
            np.testing.assert_array_equal(g[y, x], ref[min(8 * y + 4, 49), min(8 * x + 4, 69)])
    np.testing.assert_array_equal(o.guidance(ref, 70, 50, 1), ref)


def test_range_from_previous_closed_forms():
    """A previous layer at exactly plane k everywhere: the index map is k (the midpoints below it), so every pixel
    searches [k - margin, k + 1); at the near end hi clamps to n - 1; at plane 0 lo clamps to 0."""
    pl = planes(32)
    for k, lo, hi in ((10, 9, 11), (0, 0, 1), (31, 30, 31), (30, 29, 31)):
        prev = np.full((6, 8), pl[k], np.float32)
        start, count, mlo, mhi = o.range_from_previous(pl, prev, None, 16, 12)
        assert (start == lo).all() and (count == hi - lo).all(), (k, start[0, 0], count[0, 0])
        assert (mlo, mhi) == (lo, hi)
    # a step: the 4-wide window (-1..+2) around the edge spans both planes
    prev = np.full((4, 8), pl[5], np.float32)
    prev[:, 4:] = pl[20]
    start, count, _, _ = o.range_from_previous(pl, prev, None, 8, 4)
    assert start[0, 2] == 4 and start[0, 2] + count[0, 2] == 21      # x = 2 sees x = 1..4
    assert start[0, 0] == 4 and count[0, 0] == 2                      # x = 0 sees x = 0..2
    assert start[0, 4] == 4 and start[0, 5] == 19 and start[0, 7] == 19   # x = 4 sees 3..6; x = 5 sees 4..7
    # skipped pixels do not count: hide the far half's left column
    skip = np.zeros((4, 8), np.uint8)
    skip[:, 4] = 0xFF
    start, count, _, _ = o.range_from_previous(pl, prev, skip, 8, 4)
    assert start[0, 2] == 4 and count[0, 2] == 2                      # x = 2 now sees x = 1, 2, 3 only


def test_range_from_previous_resampling():
    """(int)(x * (pw-1)/(w-1)) picks the previous pixel: a previous layer whose columns sit on distinct planes."""
    pl = planes(64)
    pw, ph, w, h = 5, 3, 9, 5
    prev = np.empty((ph, pw), np.float32)
    for x in range(pw):
        prev[:, x] = pl[10 * x + 5]
    start, count, _, _ = o.range_from_previous(pl, prev, None, w, h)
    fx = np.float32(pw - 1) / np.float32(w - 1)
    for x in range(w):
        px = int(np.float32(x) * fx)
        win = [10 * min(max(px + k, 0), pw - 1) + 5 for k in (-1, 0, 1, 2)]
        assert start[0, x] == min(win) - 1 and start[0, x] + count[0, x] == max(win) + 1


def test_layer_depth_first_minimum():
    pl = planes(16)
    start = np.array([[0, 3, 10]], np.uint16)
    count = np.array([[3, 2, 5]], np.uint16)
    agg = np.array([7, 5, 5, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9,
                    4, 4, 4, 4, 1, 1, 1, 1], np.uint16)               # caps 8, 8, 8
    got = o.layer_depth(pl, 3, 1, 8, start, count, agg)
    np.testing.assert_array_equal(got[0], [pl[1], pl[3], pl[14]])


def test_oracle_skip_mask_and_generator():
    """The oracle's own MT19937: the C++ standard's 10000th output of the default seed; its skip mask equals the
    numpy-generator model of the tiling."""
    from test_abi import _skip_mask_model

    assert o.mt19937(5489, 10000) == 4123659995
    assert o.mt19937(7, 625) == int(np.random.RandomState(7).randint(0, 2 ** 32, size=625, dtype=np.uint32)[-1])
    for w, h in ((65, 49), (130, 98), (1, 1), (200, 150)):
        np.testing.assert_array_equal(o.skip_mask(2, w, h), _skip_mask_model(w, h))
        assert (o.skip_mask(0, w, h) == 0).all()


def test_memory_estimate_closed_form():
    """MemoryEstimate of a 2 x 1 layer: caps 8 and 16 (one skipped), MaxHi 9 -> 16, 100 image bytes."""
    est = o.memory_estimate(2, 1, [0, 0], [5, 9], [0, 0xFF], 8, 9, 100)
    cost = (8 + 8 * 3) + (8 + 16 * 2)
    assert est == cost + 8 + 20 * 8 + 2 * ((4 * 16 + 5) * 8 + 16 + 16) + 100 + 16
