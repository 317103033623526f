"""Pins of the CPU oracle (oracle/lux_oracle.c) against independent models, brute force and closed forms."""
import numpy as np
import pytest

from oracle import oracle as o
import sgm_model as M

SHAPES = [(W, H, T) for T in (2, 4, 8) for W in (1, 2, 3, T - 1, T, T + 1, 2 * T + 3) for H in (1, 2, T, T + 1, 2 * T + 1)
          if W >= 1 and H >= 1]


def _layer_uniform(raw, T, p1, p2, xsw=False, keep_pass1=False, align=8):
    H, W, cap = raw.shape
    start = np.zeros(W * H, np.uint16)
    count = np.full(W * H, cap, np.uint16)
    agg, agg1, wta = o.layer(raw.reshape(-1), W, H, start, count, T=T, align=align, p1=p1, p2=p2, xsw=xsw,
                             keep_pass1=keep_pass1)
    agg = agg.reshape(H, W, cap)
    return (agg, None if agg1 is None else agg1.reshape(H, W, cap), wta)


@pytest.mark.parametrize("W,H,T", SHAPES)
def test_layer_matches_label_model(W, H, T):
    rng = np.random.default_rng(W * 1000 + H * 10 + T)
    cap = 16
    raw = rng.integers(0, 49, (H, W, cap)).astype(np.uint8)
    for p1, p2, xsw in ((10, 120, False), (3, 7, True), (2500, 3000, False)):
        want, want1 = M.label_eval(raw, T, p1, p2, xsw=xsw, return_pass1=True)
        agg, agg1, wta = _layer_uniform(raw, T, p1, p2, xsw=xsw, keep_pass1=True)
        np.testing.assert_array_equal(agg, want)
        np.testing.assert_array_equal(agg1, want1)
        a = want.astype(np.int64)
        first_min = np.argmax(a == a.min(axis=2, keepdims=True), axis=2)
        np.testing.assert_array_equal(wta, first_min)


def test_layer_saturates_at_65535():
    # one zero-cost lane and penalties at the u16 limit: the other lanes grow by 255 per step along a
    # path until the saturating adds clamp them (DenseStereoSgm.cs:291-296)
    rng = np.random.default_rng(7)
    raw = np.full((2, 300, 8), 255, np.uint8)
    raw[:, :, 0] = 0
    raw[:, 290:] = rng.integers(0, 256, (2, 10, 8)).astype(np.uint8)
    want = M.label_eval(raw, 4, 65535, 65535)
    agg, _, _ = _layer_uniform(raw, 4, 65535, 65535)
    assert (want == 65535).any()
    np.testing.assert_array_equal(agg, want)


@pytest.mark.parametrize("W,H,T", [(13, 9, 4), (8, 8, 4), (17, 5, 8), (1, 6, 2)])
def test_invariants_constant_raw_and_zero_penalties(W, H, T):
    _, visits = M.label_eval(np.zeros((H, W, 8), np.uint8), T, 1, 1, return_visits=True)
    nvis = np.zeros((H, W), np.int64)
    for (d, x, y), n in visits.items():
        nvis[y, x] += n
    # constant raw c: every path cost equals c, so Agg = 4 * visits * c (8c, 12c at double visits)
    c = 9
    agg, _, _ = _layer_uniform(np.full((H, W, 8), c, np.uint8), T, 10, 120)
    np.testing.assert_array_equal(agg, (4 * c * nvis)[:, :, None].repeat(8, 2))
    # P1 = P2 = 0: m = minPrev, every path cost equals raw
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 60, (H, W, 8)).astype(np.uint8)
    agg, _, _ = _layer_uniform(raw, T, 0, 0)
    np.testing.assert_array_equal(agg, raw.astype(np.int64) * 4 * nvis[:, :, None])
    # D = 1 (align 1): the neighbour lanes are guards, every path cost equals raw
    raw1 = rng.integers(0, 60, (H, W, 1)).astype(np.uint8)
    agg, _, _ = _layer_uniform(raw1, T, 10, 120, align=1)
    np.testing.assert_array_equal(agg, raw1.astype(np.int64) * 4 * nvis[:, :, None])
    assert nvis.min() >= 2 and nvis.max() <= (4 if W == 1 else 3)


def test_double_visit_counts_640x480_t4():
    # every cell is visited >= once per pass; 640x480 with T=4 double-visits 120 cells per pass
    vis = {}
    for d, x, y, cur in M.schedule(640, 480, 4):
        vis[(d, x, y)] = vis.get((d, x, y), 0) + 1
    assert len(vis) == 2 * 640 * 480
    assert sum(1 for k, n in vis.items() if n == 2 and k[0] == 1) == 120
    assert all(k[1] == 639 for k, n in vis.items() if n == 2 and k[0] == 1)
    assert all(k[1] == 0 for k, n in vis.items() if n == 2 and k[0] == -1)


def test_vector_form_equals_scalar_form_variable_ranges():
    rng = np.random.default_rng(3)
    W, H, n = 23, 11, 40
    start = rng.integers(0, 20, W * H).astype(np.uint16)
    count = rng.integers(1, 21, W * H).astype(np.uint16)
    caps = ((count + 7) // 8) * 8
    raw = rng.integers(0, 256, int(caps.sum())).astype(np.uint8)
    skip = (rng.random(W * H) < 0.2).astype(np.uint8)
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
    outs = []
    for force in (0, 1):
        o.lib().lo_set_force_scalar(force)
        try:
            outs.append(o.layer(raw, W, H, start, count, T=4, p1=1, skip=skip, guidance=guid,
                                gw=(0.12, 0.03, 0.03, 0.0), p2_scale=500.0, adaptive=True, keep_pass1=True))
        finally:
            o.lib().lo_set_force_scalar(0)
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    assert n >= 20


RANGE_CASES = [(9, 7, 4, 48, 20, 8, True), (13, 5, 2, 40, 13, 8, False), (6, 11, 4, 64, 24, 8, True),
               (10, 10, 8, 32, 9, 2, False), (5, 3, 2, 16, 5, 2, True), (1, 6, 2, 24, 7, 8, False)]


@pytest.mark.parametrize("W,H,T,planes,cmax,align,smooth", RANGE_CASES)
def test_variable_ranges_match_frame_model(W, H, T, planes, cmax, align, smooth):
    """Per-pixel ranges whose padding stays inside the segments: the oracle's literal line buffer (LIdx with
    MinLo, stale cleanup, Rd) equals the frame model (a neighbour's plane d is its stored value, else 2000)."""
    rng = np.random.default_rng(W * 1000 + H * 10 + T)
    start, count = M.random_ranges(rng, W, H, planes, cmax, align, smooth)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    minlo = int(start.min())
    maxhi = int(((start.astype(np.int64) + count).max() + align - 1) // align * align)
    assert (start.astype(np.int64) - minlo + caps <= maxhi).all()
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 40, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    packed = np.concatenate(raws) if raws else np.zeros(0, np.uint8)
    for p1, p2 in ((8, 96), (0, 0), (3000, 3500)):
        agg, agg1, wta = o.layer(packed, W, H, start, count, T=T, align=align, p1=p1, p2=p2, keep_pass1=True)
        ragg, ragg1, rwta = M.label_eval_ranges(raws, start, count, align, T, p1, p2)
        np.testing.assert_array_equal(agg1, ragg1)
        np.testing.assert_array_equal(agg, ragg)
        np.testing.assert_array_equal(wta, rwta)


def test_variable_range_padding_overflow_differs_from_frame_model():
    """The case the device path refuses: a range's padding (count 3 -> cap 8) runs past MaxHi, so the
    reference writes into the next path's segment and the frame model no longer describes it."""
    W, H, T, align = 4, 3, 2, 8
    start = np.array([[0, 0, 0, 0], [0, 5, 5, 0], [0, 5, 5, 0]], np.uint16)
    count = np.array([[8, 8, 8, 8], [8, 3, 3, 8], [8, 3, 3, 8]], np.uint16)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    assert int((start.astype(np.int64) + caps).max()) > 8  # MaxHi = 8, MinLo = 0
    rng = np.random.default_rng(1)
    raws = [rng.integers(0, 9, int(c)).astype(np.uint8) for c in caps.flat]
    for r, c in zip(raws, count.flat):
        r[int(c):] = 255
    agg, _, _ = o.layer(np.concatenate(raws), W, H, start, count, T=T, align=align, p1=8, p2=96)
    ragg, _, _ = M.label_eval_ranges(raws, start, count, align, T, 8, 96)
    assert not np.array_equal(agg, ragg)


def _f32(bits):
    return np.array([bits], np.uint32).view(np.float32)[0]


def _p2_numpy(p2base, g, q, gw):
    """float32 twin of the fast-exp2 P2, one operation at a time (no FMA)."""
    f32 = np.float32
    d = [(f32(g[k]) - f32(q[k])) * f32(gw[k]) for k in range(3)]
    S = (abs(d[0]) + abs(d[2])) + abs(d[1])
    e = max(min(-S, f32(128.0)), f32(-126.0))
    ei = int(e) + (-1 if e < 0 else 0)
    f = f32(e - f32(ei))
    poly = ((_f32(0x3d9fcb52) * f + _f32(0x3e677e26)) * f + _f32(0x3f322226)) * f + _f32(0x3f7ffb19)
    bits = int(np.array([poly], np.float32).view(np.uint32)[0]) + (ei << 23)
    E = np.array([bits & 0xFFFFFFFF], np.uint32).view(np.float32)[0]
    return int(f32(p2base) * E)


def test_p2_fast_exp2():
    import ctypes
    def call(p2base, g, q, gw):
        G = (ctypes.c_float * 4)(*g); Q = (ctypes.c_float * 4)(*q); GW = (ctypes.c_float * 4)(*gw)
        return o.lib().lo_p2_adaptive(ctypes.c_float(p2base), ctypes.addressof(G), ctypes.addressof(Q), ctypes.addressof(GW))
    # closed form at S = 0: P2 = trunc(p2Base * 0.99992520) -> P1=10, P2Scale=12 gives 119, not 120
    assert call(120.0, [5, 6, 7, 8], [5, 6, 7, 8], [1, 1, 1, 0]) == 119
    assert call(np.float32(10 * 12.001), [0] * 4, [0] * 4, [0] * 4) == 120
    c = _f32(0x3ef6384f)
    gw = [c / 6, c / 16, c / 16, 0.0]
    rng = np.random.default_rng(0)
    for _ in range(200):
        g = rng.integers(0, 256, 4).astype(np.float32)
        q = rng.integers(0, 256, 4).astype(np.float32)
        assert call(500.0, list(g), list(q), gw) == _p2_numpy(500.0, g, q, gw)
    # S exactly an integer: e = -2 gives ei = -3, f = 1 (the reference's truncation rule)
    assert call(1000.0, [2, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]) == _p2_numpy(1000.0, [2, 0, 0, 0], [0] * 4, [1, 0, 0, 0])
    # exp2 closed form at integer exponents, up to the cubic's accuracy
    for S in (1, 3, 5):
        assert abs(call(4096.0, [S, 0, 0, 0], [0] * 4, [1, 0, 0, 0]) - 4096.0 * 2.0 ** -S) <= 4096.0 * 2.0 ** -S * 1e-3 + 1


@pytest.mark.parametrize("radius,dtype", [(2, np.uint8), (3, np.uint16), (1, np.uint8)])
def test_census_bruteforce(radius, dtype):
    rng = np.random.default_rng(radius)
    img = rng.integers(0, 1024 if dtype == np.uint16 else 256, (9, 12)).astype(dtype)
    np.testing.assert_array_equal(o.census(img, radius), M.census_bruteforce(img, radius))
    img[:] = 5  # flat image: no neighbour is below the centre
    assert not o.census(img, radius).any()


def test_cost_census_bruteforce_and_threads():
    rng = np.random.default_rng(4)
    L = rng.integers(0, 256, (7, 15)).astype(np.uint8)
    R = rng.integers(0, 256, (7, 15)).astype(np.uint8)
    cl, cr = o.census(L, 2), o.census(R, 2)
    raw = o.cost_census(cl, cr, 10, 16)
    np.testing.assert_array_equal(raw, M.cost_bruteforce(cl, cr, 10, 16))
    np.testing.assert_array_equal(raw, o.cost_census(cl, cr, 10, 16, threads=4))
    assert (raw[:, :, 10:] == 255).all()
    # a plane left of the image repeats the last in-image plane
    np.testing.assert_array_equal(raw[:, 3, 4:10], raw[:, 3, 3:4].repeat(6, 1))


def test_cost_ad_definition():
    rng = np.random.default_rng(5)
    L = rng.integers(0, 1024, (6, 11)).astype(np.uint16)
    R = rng.integers(0, 1024, (6, 11)).astype(np.uint16)
    raw = o.cost_ad(L, R, 10, 7, 8, adcap=28)
    H, W = L.shape
    for y in range(H):
        for x in range(W):
            for d in range(7):
                xr = max(x - d, 0)
                s = 0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        yy = min(max(y + dy, 0), H - 1)
                        a = int(L[yy, min(max(x + dx, 0), W - 1)])
                        b = int(R[yy, min(max(xr + dx, 0), W - 1)])
                        s += min(abs(a - b) >> 2, 28)
                assert raw[y, x, d] == s
    assert (raw[:, :, 7] == 255).all() and raw[:, :, :7].max() <= 252


@pytest.mark.parametrize("shape,dtype,maxval", [((9, 13), np.uint8, 255), ((1, 7), np.uint16, 1023), ((6, 1), np.uint8, 255)])
def test_box3_bruteforce(shape, dtype, maxval):
    rng = np.random.default_rng(8)
    img = rng.integers(0, maxval + 1, shape).astype(dtype)
    np.testing.assert_array_equal(o.box3(img, maxval), M.box3_bruteforce(img, maxval))


def test_wta_subpixel_lr_closed_forms():
    # hand-built triples: l=10, c=4, r=6 -> den 8, offset 0.5*(10-6)/8 = 0.25
    agg = np.full((1, 3, 8), 50, np.uint16)
    agg[0, 0, 2:5] = (10, 4, 6)
    agg[0, 1, 0] = 3                      # minimum at the range start: no offset
    agg[0, 2, 7] = 1                      # minimum at the range end: no offset
    agg[0, 2, 3] = 1                      # ...but the first minimum wins (strict <)
    d, s, v = o.wta_sub_lr(agg, 8, tol=1)
    assert list(d[0]) == [3, 0, 3]
    assert s[0, 0] == np.float32(3.25) and s[0, 1] == 0.0
    # x=2 picked d=3 -> x-d < 0 -> invalid; x=1 picked 0 -> right pixel 1
    assert v[0, 2] == 0
    rng = np.random.default_rng(9)
    agg = rng.integers(0, 300, (4, 20, 16)).astype(np.uint16)
    got = o.wta_sub_lr(agg, 13, tol=1)
    want = M.wta_sub_lr_model(agg, 13, tol=1)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_known_shift_recovered():
    # the reference's known-answer idiom (StereoTests.cs:41-58): a textured pair shifted by a known disparity
    rng = np.random.default_rng(11)
    L = rng.integers(0, 256, (40, 64)).astype(np.uint8)
    R = np.zeros_like(L)
    R[:, :-5] = L[:, 5:]
    R[:, -5:] = L[:, -5:]
    out = o.rectified(L, R, 16, radius=2, bits=8, p1=10, p2=120, T=4)
    assert (out["disp"][:, 8:] == 5).mean() > 0.97
    assert out["valid"][:, 8:].mean() > 0.9
    # same input, same bytes (StereoTests.cs:130-138)
    out2 = o.rectified(L, R, 16, radius=2, bits=8, p1=10, p2=120, T=4, threads=3)
    for k in ("disp", "sub", "valid"):
        np.testing.assert_array_equal(out[k], out2[k])


def test_downsample2_bruteforce():
    rng = np.random.default_rng(4)
    for H, W in ((5, 7), (6, 8), (1, 3), (9, 1)):
        img = rng.integers(0, 1024, (H, W)).astype(np.uint16)
        out = o.downsample2(img)
        for y in range((H + 1) // 2):
            for x in range((W + 1) // 2):
                ys, xs = [min(2 * y, H - 1), min(2 * y + 1, H - 1)], [min(2 * x, W - 1), min(2 * x + 1, W - 1)]
                s = sum(int(img[a, b]) for a in ys for b in xs)
                assert out[y, x] == (s + 2) >> 2


def test_cost_census_ranges_bruteforce():
    rng = np.random.default_rng(5)
    H, W, n = 4, 9, 6
    img_l = rng.integers(0, 256, (H, W)).astype(np.uint8)
    img_r = rng.integers(0, 256, (H, W)).astype(np.uint8)
    cl, cr = o.census(img_l, 2), o.census(img_r, 2)
    start = rng.integers(0, 12, (H, W))
    raw = o.cost_census_ranges(cl, cr, start, n)
    for y in range(H):
        for x in range(W):
            for j in range(n):
                want = bin(int(cl[y, x]) ^ int(cr[y, max(x - (int(start[y, x]) + j), 0)])).count("1")
                assert raw[y, x, j] == want


def test_pyramid_recovers_a_known_shift():
    """A pure shift by d at full resolution is d/2 one level down: every level must find it (away from the
    left border, where the right image has no match)."""
    rng = np.random.default_rng(6)
    H, W, d = 48, 96, 12
    base = rng.integers(0, 1024, (H, W + d)).astype(np.uint16)
    left = o.box3(base, 1023)[:, :W].copy()        # left(x) = right(x - d)
    right = o.box3(base, 1023)[:, d:].copy()
    disp, sub = o.pyramid(left, right, 32, [4], radius=2, bits=10, T=8)
    inner = disp[4:-4, 24:-4]
    assert (inner == d).mean() > 0.98
    assert np.abs(sub[4:-4, 24:-4] - d).max() < 1.0
