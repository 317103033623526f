"""Parity of the gfx950 path with the CPU oracle (bit-exact: integer disparities, aggregates, validity;
subpixel compared bit for bit too, within the stated 1e-4 px tolerance a fortiori)."""
import numpy as np
import pytest

from lux_amd import scene, sgm
from oracle import oracle as o

pytestmark = pytest.mark.gpu
SUB_TOL = 1e-4  # px


@pytest.fixture(scope="module", params=["sweeps", "bands"])
def ctx(request):
    c = sgm.Context(0)
    c.set_option(sgm.OPT_FUSED_BANDS, int(request.param == "bands"))
    yield c
    c.close()


def _layer_case(ctx, W, H, planes, T, p1, p2, align=8, xsw=False, seed=0, hi=49):
    rng = np.random.default_rng(seed)
    cap = ((planes + align - 1) // align) * align
    raw = rng.integers(0, hi, (H, W, cap)).astype(np.uint8)
    raw[:, :, planes:] = 255
    print(f"layer W={W} H={H} planes={planes} cap={cap} T={T} p1={p1} p2={p2} xsw={xsw}", flush=True)
    L = sgm.DenseLayer(ctx, W, H, planes, align=align, p1=p1, p2=p2, tile=T, xsw=xsw)
    L.Raw = raw
    L.KeepPass1 = True
    L.compute()
    start = np.zeros(W * H, np.uint16)
    count = np.full(W * H, planes, np.uint16)
    agg, agg1, wta = o.layer(raw.reshape(-1), W, H, start, count, T=T, align=align, p1=p1, p2=p2, xsw=xsw,
                             keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1.reshape(H, W, cap))
    np.testing.assert_array_equal(L.Agg, agg.reshape(H, W, cap))
    np.testing.assert_array_equal(L.Wta, wta)


@pytest.mark.parametrize("W,H,T", [(1, 1, 2), (1, 7, 2), (2, 5, 4), (3, 3, 2), (7, 9, 4), (8, 8, 4), (9, 13, 4),
                                   (16, 9, 8), (17, 17, 8), (33, 12, 16), (70, 40, 64), (130, 70, 64)])
def test_layer_shapes(ctx, W, H, T):
    _layer_case(ctx, W, H, 16, T, 10, 120, seed=W * 100 + H)
    _layer_case(ctx, W, H, 16, T, 10, 120, seed=W * 100 + H + 1, xsw=True)


@pytest.mark.parametrize("planes,align", [(8, 8), (13, 8), (64, 8), (96, 8), (128, 8), (192, 8), (256, 8),
                                          (320, 8), (512, 8), (6, 2), (66, 2), (1000, 8)])
def test_layer_plane_counts(ctx, planes, align):
    _layer_case(ctx, 21, 11, planes, 4, 8, 96, align=align, seed=planes)


def test_layer_penalties(ctx):
    _layer_case(ctx, 19, 14, 24, 4, 0, 0, seed=1)
    _layer_case(ctx, 19, 14, 24, 4, 2500, 3000, seed=2)   # penalties past the 2000 guards
    _layer_case(ctx, 40, 3, 8, 4, 65535, 65535, seed=3, hi=256)  # saturating adds


def _rect_case(ctx, W, H, D, *, bits, cost, radius=3, T=64, p1=8, p2=96, seed=0, planes=6, sigma=4.0, xsw=False):
    left, right, _ = scene.make_pair(W, H, max(D - 1, 1), planes=planes, sigma=sigma, seed=seed, bits=bits)
    print(f"rectified W={W} H={H} D={D} bits={bits} cost={cost} r={radius} T={T}", flush=True)
    got = sgm.rectified(ctx, left, right, D, want_agg=True, bits=bits, cost=cost, radius=radius, p1=p1, p2=p2,
                        tile=T, xsw=xsw)
    want = o.rectified(left, right, D, cost=cost, radius=radius, bits=bits, p1=p1, p2=p2, T=T, xsw=xsw, want_agg=True)
    np.testing.assert_array_equal(got["agg"], want["agg"])
    np.testing.assert_array_equal(got["disp"], want["disp"])
    np.testing.assert_array_equal(got["valid"], want["valid"])
    assert np.abs(got["sub"] - want["sub"]).max() <= SUB_TOL
    np.testing.assert_array_equal(got["sub"], want["sub"])
    return got, want


@pytest.mark.parametrize("bits,cost,radius", [(8, "census", 2), (10, "census", 3), (8, "ad", 0), (10, "ad", 0),
                                              (12, "census", 1)])
def test_rectified_front_ends(ctx, bits, cost, radius):
    _rect_case(ctx, 97, 53, 32, bits=bits, cost=cost, radius=radius or 3, T=8, seed=bits)


@pytest.mark.parametrize("D,p1,p2", [(32, 8, 96), (30, 8, 96), (32, 20, 300), (256, 8, 207), (256, 8, 208)])
def test_rectified_census_byte_bound(ctx, D, p1, p2):
    """Census 7x7 (costs up to 48): P2 <= 207 with no padding planes runs with byte path volumes, a padded plane count
    (30 of cap 32) or P2 >= 208 with the u16 aggregate; every one equals the oracle."""
    _rect_case(ctx, 75, 41, D, bits=10, cost="census", radius=3, T=16, p1=p1, p2=p2, seed=D + p2)


def test_rectified_config0(ctx):
    # 640x480 uint8, census 5x5, D=64, P1=10, P2=120, L/R tolerance 1 (seeded planar scene, 4 planes, sigma 2)
    got, want = _rect_case(ctx, 640, 480, 64, bits=8, cost="census", radius=2, T=64, p1=10, p2=120, planes=4, sigma=2.0)
    assert got["valid"].mean() > 0.8


def test_rectified_mid_size_u16(ctx):
    _rect_case(ctx, 1040, 780, 96, bits=10, cost="census", radius=3, T=64, p1=8, p2=96, seed=5)


def test_full_module_properties(ctx):
    # 4160x3120 uint16 10-bit, census 7x7, D=256: deterministic, accurate against the generator's field
    W, H, D = 4160, 3120, 256
    left, right, truth = scene.make_pair(W, H, D - 1, planes=6, sigma=4.0, seed=1, bits=10)
    print("full module 4160x3120 D=256", flush=True)
    a = sgm.rectified(ctx, left, right, D, bits=10, cost="census", radius=3, p1=8, p2=96, tile=64)
    b = sgm.rectified(ctx, left, right, D, bits=10, cost="census", radius=3, p1=8, p2=96, tile=64)
    for k in ("disp", "sub", "valid"):
        np.testing.assert_array_equal(a[k], b[k])
    v = a["valid"] > 0
    assert v.mean() > 0.8
    err = np.abs(a["disp"].astype(np.int64) - truth)
    assert (err[v] <= 1).mean() > 0.95
    assert np.abs(a["sub"] - a["disp"]).max() <= 0.5


# ------------------------------------------------------------------ per-pixel plane ranges
import sgm_model as M  # noqa: E402

RANGE_LAYERS = [(9, 7, 4, 48, 20, 8, True), (13, 5, 2, 40, 13, 8, False), (33, 12, 16, 64, 24, 8, True),
                (10, 10, 8, 32, 9, 2, False), (1, 6, 2, 24, 7, 8, False), (70, 40, 64, 128, 40, 8, True),
                (130, 70, 64, 256, 128, 8, True), (64, 33, 8, 200, 99, 2, False), (40, 9, 4, 96, 1, 8, True),
                (130, 70, 64, 300, 200, 8, True), (70, 40, 16, 512, 400, 8, False), (64, 33, 8, 300, 250, 2, True),
                (33, 21, 4, 520, 512, 8, False),
                # taller than wide and square, caps above 32: the unsegmented pass and both branches of its
                # longest-first diagonal order
                (40, 70, 8, 128, 40, 8, True), (48, 48, 8, 128, 40, 8, False)]


@pytest.mark.parametrize("W,H,T,planes,cmax,align,smooth", RANGE_LAYERS)
def test_layer_ranges(W, H, T, planes, cmax, align, smooth):
    rng = np.random.default_rng(W * 1000 + H * 10 + T + cmax)
    start, count = M.random_ranges(rng, W, H, planes, cmax, align, smooth)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    packed = np.concatenate(raws)
    print(f"ranges W={W} H={H} T={T} planes={planes} cmax={cmax} align={align}", flush=True)
    with sgm.Context(0) as c:
        for p1, p2 in ((8, 96), (2500, 3000)):
            L = sgm.DenseLayer(c, W, H, planes, align=align, p1=p1, p2=p2, tile=T).set_ranges(start, count)
            L.Raw = packed
            L.KeepPass1 = True
            L.compute()
            agg, agg1, wta = o.layer(packed, W, H, start, count, T=T, align=align, p1=p1, p2=p2, keep_pass1=True)
            np.testing.assert_array_equal(L.AggPass1, agg1)
            np.testing.assert_array_equal(L.Agg, agg)
            np.testing.assert_array_equal(L.Wta, wta)


def _ranged_raw(rng, count, align):
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raws = []
    for i in range(count.size):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    return np.concatenate(raws)


def _spills(start, count, align):
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    maxhi = -(-int((start.astype(np.int64) + count).max()) // align) * align
    return int((start.astype(np.int64) - int(start.min()) + caps > maxhi).sum())


def test_layer_ranges_padding_overflow_small():
    # MinLo 0 and ranges [5, 8) with cap 8: 5 - 0 + 8 > 8, the padding runs into the next segment
    start = np.array([[0, 0, 0, 0], [0, 5, 5, 0], [0, 5, 5, 0]], np.uint16)
    count = np.array([[8, 8, 8, 8], [8, 3, 3, 8], [8, 3, 3, 8]], np.uint16)
    raw = _ranged_raw(np.random.default_rng(3), count, 8)
    with sgm.Context(0) as c:
        L = sgm.DenseLayer(c, 4, 3, 16, align=8, p1=8, p2=96, tile=2).set_ranges(start, count)
        L.Raw = raw
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(raw, 4, 3, start, count, T=2, align=8, p1=8, p2=96, keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1)
    np.testing.assert_array_equal(L.Agg, agg)
    np.testing.assert_array_equal(L.Wta, wta)


# W, H, T, planes, cmax, align, adaptive: padding past the segments, odd tiles, caps above 128 (4 planes a lane)
SPILL_LAYERS = [(9, 7, 4, 40, 13, 8, False), (13, 11, 3, 24, 11, 8, False), (16, 12, 5, 30, 9, 2, True),
                (70, 40, 64, 96, 45, 8, True), (130, 70, 16, 160, 77, 8, False), (200, 90, 8, 300, 130, 8, True),
                (41, 33, 7, 64, 61, 8, False), (300, 120, 64, 120, 40, 8, True), (60, 30, 8, 800, 700, 8, False),
                (33, 20, 5, 400, 290, 8, True)]


@pytest.mark.parametrize("W,H,T,planes,cmax,align,adaptive", SPILL_LAYERS)
def test_layer_ranges_spilling_padding(W, H, T, planes, cmax, align, adaptive):
    _spill_case(W, H, T, planes, cmax, align, adaptive)


@pytest.mark.parametrize("local", ["0", "1"])
def test_layer_spilling_padding_on_one_xcd_and_on_all(monkeypatch, local):
    """The line-buffer-literal pass runs on one XCD (its words through that XCD's L2) when every row fits there at
    once, else on all of them (words coherent across XCDs); LUX_SGM_LIT_LOCAL forces either. Both equal the oracle
    (planes up to 128, 256 and 1024 a cell: two, four and the LDS form)."""
    monkeypatch.setenv("LUX_SGM_LIT_LOCAL", local)
    for case in (SPILL_LAYERS[3], SPILL_LAYERS[5], SPILL_LAYERS[8]):
        _spill_case(*case)


def _spill_case(W, H, T, planes, cmax, align, adaptive):
    rng = np.random.default_rng(W * 1000 + H + T)
    start, count = M.spill_ranges(rng, W, H, planes, cmax, align)
    assert _spills(start, count, align) > 0
    raw = _ranged_raw(rng, count, align)
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8) if adaptive else None
    print(f"spill W={W} H={H} T={T} planes={planes} cmax={cmax} align={align} adaptive={adaptive}", flush=True)
    kw = dict(T=T, align=align, keep_pass1=True)
    if adaptive:
        kw.update(p1=1, guidance=guid, gw=L16_WEIGHTS, p2_scale=500.0, adaptive=True)
    else:
        kw.update(p1=8, p2=96)
    with sgm.Context(0) as c:
        L = sgm.DenseLayer(c, W, H, planes, align=align, p1=kw["p1"], p2=kw.get("p2", 0), tile=T)
        L.set_ranges(start, count)
        if adaptive:
            L.set_guidance(guid, L16_WEIGHTS, 500.0)
        L.Raw = raw
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(raw, W, H, start, count, **kw)
    np.testing.assert_array_equal(L.AggPass1, agg1)
    np.testing.assert_array_equal(L.Agg, agg)
    np.testing.assert_array_equal(L.Wta, wta)


@pytest.mark.parametrize("W,H,T,xsw,skipped", [(37, 29, 4, True, False), (65, 50, 8, False, True), (50, 37, 5, True, True)])
def test_layer_spilling_padding_xsw_and_skip(W, H, T, xsw, skipped):
    """Spilling padding with the LUX_SGM_XSW clamp (no second visit in pass +1) and with a mode-2 skip mask."""
    planes, cmax, align = 48, 21, 8
    rng = np.random.default_rng(W * H + T)
    start, count = M.spill_ranges(rng, W, H, planes, cmax, align)
    assert _spills(start, count, align) > 0
    raw = _ranged_raw(rng, count, align)
    skip = sgm.skip_mask(2, W, H) if skipped else None
    print(f"spill xsw={xsw} skip={skipped} W={W} H={H} T={T}", flush=True)
    with sgm.Context(0) as c:
        L = sgm.DenseLayer(c, W, H, planes, align=align, p1=8, p2=96, tile=T, xsw=xsw).set_ranges(start, count)
        if skipped:
            L.Skip = skip
        L.Raw = raw
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(raw, W, H, start, count, T=T, align=align, p1=8, p2=96, xsw=xsw, skip=skip,
                             keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1)
    np.testing.assert_array_equal(L.Agg, agg)
    np.testing.assert_array_equal(L.Wta, wta)


@pytest.mark.parametrize("W,H,T,planes", [(9, 13, 3, 16), (33, 12, 5, 64), (70, 40, 63, 24), (17, 17, 9, 200)])
def test_layer_odd_tiles(ctx, W, H, T, planes):
    _layer_case(ctx, W, H, planes, T, 10, 120, seed=W + H + T)
    _layer_case(ctx, W, H, planes, T, 10, 120, seed=W + H + T + 1, xsw=True)


def test_layer_literal_path_forced():
    """The line-buffer-literal path on layers the path sweeps also take (LUX_SGM_OPT_LITERAL): the same results."""
    with sgm.Context(0) as c:
        c.set_option(sgm.OPT_LITERAL, 1)
        for W, H, T, planes in ((7, 9, 4, 16), (130, 70, 64, 96), (64, 33, 8, 256)):
            _layer_case(c, W, H, planes, T, 8, 96, seed=W * H)
        rng = np.random.default_rng(8)
        W, H, T, planes, cmax, align = 90, 50, 16, 128, 60, 8
        start, count = M.random_ranges(rng, W, H, planes, cmax, align, True)
        raw = _ranged_raw(rng, count, align)
        L = sgm.DenseLayer(c, W, H, planes, align=align, p1=8, p2=96, tile=T).set_ranges(start, count)
        L.Raw = raw
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(raw, W, H, start, count, T=T, align=align, p1=8, p2=96, keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1)
    np.testing.assert_array_equal(L.Agg, agg)
    np.testing.assert_array_equal(L.Wta, wta)


# ------------------------------------------------------------------ coarse-to-fine requests
@pytest.mark.parametrize("W,H,D,radii,bits,fg,T", [(160, 120, 64, (8, 4), 10, 0.1, 16), (130, 66, 64, (8,), 8, 0.0, 8),
                                                  (257, 131, 128, (16, 8), 10, 0.1, 64),
                                                  (96, 80, 32, (4, 4, 4), 10, 0.0, 4)])
def test_pyramid_matches_oracle(W, H, D, radii, bits, fg, T):
    left, right, _ = scene.make_pair(W, H, D - 1, planes=4, sigma=2.0, seed=W + H, bits=bits, fg_frac=fg)
    print(f"pyramid W={W} H={H} D={D} radii={radii} bits={bits} T={T}", flush=True)
    with sgm.Context(0) as c:
        r = sgm.pyramid(c, left, right, D, radii, bits=bits, radius=3, p1=8, p2=96, tile=T)
    disp, sub = o.pyramid(left, right, D, list(radii), radius=3, bits=bits, p1=8, p2=96, T=T)
    np.testing.assert_array_equal(r["disp"], disp)
    np.testing.assert_allclose(r["sub"], sub, rtol=0, atol=SUB_TOL)
    np.testing.assert_array_equal(r["sub"], sub)


# ------------------------------------------------------------------ adaptive P2 (guidance)
C_G = float(np.array([0x3ef6384f], np.uint32).view(np.float32)[0])
L16_WEIGHTS = (C_G / 6, C_G / 16, C_G / 16, 0.0)  # StereoParams.L16Pyramid guidance weights (DenseStereo.cs:17-21)


@pytest.mark.parametrize("W,H,T,planes,p1,scale,weights", [
    (21, 11, 4, 24, 1, 500.0, L16_WEIGHTS), (33, 17, 8, 64, 8, 12.0, (0.3, 0.02, 0.05, 0.7)),
    (9, 13, 2, 16, 10, 12.0, (2.0, 2.0, 2.0, 0.0)), (70, 40, 64, 256, 3, 40.0, L16_WEIGHTS)])
def test_layer_adaptive_p2(ctx, W, H, T, planes, p1, scale, weights):
    rng = np.random.default_rng(W * H + planes)
    cap = ((planes + 7) // 8) * 8
    raw = rng.integers(0, 49, (H, W, cap)).astype(np.uint8)
    raw[:, :, planes:] = 255
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
    guid[H // 2:, : W // 2] = guid[0, 0]  # flat regions too (S = 0)
    print(f"adaptive W={W} H={H} T={T} planes={planes} p1={p1} scale={scale}", flush=True)
    L = sgm.DenseLayer(ctx, W, H, planes, p1=p1, p2=0, tile=T).set_guidance(guid, weights, scale)
    L.Raw = raw
    L.KeepPass1 = True
    L.compute()
    start = np.zeros(W * H, np.uint16)
    count = np.full(W * H, planes, np.uint16)
    agg, agg1, wta = o.layer(raw.reshape(-1), W, H, start, count, T=T, p1=p1, guidance=guid, gw=weights, p2_scale=scale,
                             adaptive=True, keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1.reshape(H, W, cap))
    np.testing.assert_array_equal(L.Agg, agg.reshape(H, W, cap))
    np.testing.assert_array_equal(L.Wta, wta)


@pytest.mark.parametrize("W,H,T,planes,cmax,align", [(40, 23, 8, 96, 40, 8), (50, 31, 8, 400, 300, 8)])
def test_layer_ranges_adaptive_p2(W, H, T, planes, cmax, align):
    rng = np.random.default_rng(11)
    start, count = M.random_ranges(rng, W, H, planes, cmax, align, True)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    packed = np.concatenate(raws)
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
    with sgm.Context(0) as c:
        L = sgm.DenseLayer(c, W, H, planes, align=align, p1=1, p2=0, tile=T).set_ranges(start, count)
        L.set_guidance(guid, L16_WEIGHTS, 500.0)
        L.Raw = packed
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(packed, W, H, start, count, T=T, align=align, p1=1, guidance=guid, gw=L16_WEIGHTS,
                             p2_scale=500.0, adaptive=True, keep_pass1=True)
    np.testing.assert_array_equal(L.AggPass1, agg1)
    np.testing.assert_array_equal(L.Agg, agg)
    np.testing.assert_array_equal(L.Wta, wta)


# ------------------------------------------------------------------ skip masks
@pytest.mark.parametrize("ranged", [False, True])
def test_layer_skip_mask(ranged):
    W, H, T, planes, align = 130, 98, 8, 48, 8
    rng = np.random.default_rng(17)
    skip = sgm.skip_mask(2, W, H)
    if ranged:
        start, count = M.random_ranges(rng, W, H, planes, 24, align, True)
    else:
        start, count = np.zeros((H, W), np.uint16), np.full((H, W), planes, np.uint16)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    packed = np.concatenate(raws)
    with sgm.Context(0) as c:
        L = sgm.DenseLayer(c, W, H, planes, align=align, p1=8, p2=96, tile=T)
        if ranged:
            L.set_ranges(start, count)
        L.Skip = skip
        L.Raw = packed if ranged else packed.reshape(H, W, -1)
        L.KeepPass1 = True
        L.compute()
    agg, agg1, wta = o.layer(packed, W, H, start, count, T=T, align=align, p1=8, p2=96, skip=skip, keep_pass1=True)
    np.testing.assert_array_equal(np.asarray(L.AggPass1).reshape(-1), agg1)
    np.testing.assert_array_equal(np.asarray(L.Agg).reshape(-1), agg)
    np.testing.assert_array_equal(L.Wta, wta)


# ------------------------------------------------------------------ Lumen dumps as the wire format
def test_layer_through_lumen_dump(tmp_path):
    from lux_amd import lumen

    W, H, T, planes, align = 37, 21, 4, 64, 8
    rng = np.random.default_rng(23)
    start, count = M.random_ranges(rng, W, H, planes, 24, align, True)
    skip = sgm.skip_mask(2, W, H)
    caps = lumen.caps_of(count.reshape(-1), align)
    raw = rng.integers(0, 49, int(caps.sum())).astype(np.uint8)
    src = str(tmp_path / "in")
    lumen.write_costvol(src, W, H, start, count, align, np.zeros(int(caps.sum()), np.uint16), raw=raw, skip=skip)
    with sgm.Context(0) as c:
        L = lumen.layer_from_dump(c, src, planes, align=align, p1=8, p2=96, tile=T)
        L.KeepPass1 = True
        L.compute()
    out = str(tmp_path / "gpu")
    lumen.dump_layer(out, L)
    d = lumen.read_costvol(src)
    agg, agg1, _ = o.layer(d["raw"], W, H, start, count, T=T, align=align, p1=8, p2=96, skip=skip, keep_pass1=True)
    ref = str(tmp_path / "ref")
    lumen.write_costvol(ref, W, H, start, count, align, agg, raw=d["raw"], skip=skip, agg_pass1=agg1)
    for suffix in ("_offsets.u32", "_costvol.bin", "_costvol_pass1.bin"):
        assert open(out + suffix, "rb").read() == open(ref + suffix, "rb").read(), suffix


# ------------------------------------------------------------------ box filter
@pytest.mark.parametrize("shape,dtype,bits", [((97, 131), np.uint8, 8), ((64, 48), np.uint16, 10), ((1, 5), np.uint8, 8),
                                              ((7, 1), np.uint16, 16), ((33, 20, 4), np.uint8, 8), ((1, 1, 4), np.uint8, 8),
                                              ((1, 5, 4), np.uint8, 8), ((2, 3, 4), np.uint8, 8), ((130, 97, 4), np.uint8, 8)])
def test_box3(shape, dtype, bits):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 1 << bits, shape).astype(dtype)
    with sgm.Context(0) as c:
        got = sgm.box3(c, img, bits)
    if img.ndim == 3:
        want = np.stack([o.box3(np.ascontiguousarray(img[:, :, k]), 255) for k in range(4)], axis=2)
    else:
        want = o.box3(img, (1 << bits) - 1)
    np.testing.assert_array_equal(got, want)


# ------------------------------------------------------------------ DenseUpsampleLayer
def _upsample_inputs(pw, ph, W, H, seed):
    rng = np.random.default_rng(seed)
    depth = rng.uniform(0.5, 40.0, (ph, pw)).astype(np.float32)
    depth[rng.random((ph, pw)) < 0.02] = 0.0          # holes: rcp(0) = inf, as the reference's inverse leaves them
    base = rng.integers(0, 256, (H // 4 + 2, W // 4 + 2, 4)).astype(np.uint8)
    guide = np.repeat(np.repeat(base, 4, 0), 4, 1)[:H, :W].copy()   # blocky, so the range weights vary widely
    guide[rng.random((H, W)) < 0.1] = rng.integers(0, 256, 4)
    conf = rng.uniform(0.0, 1.0, (ph, pw)).astype(np.float32)
    return depth, guide, conf


def _upsample_conf_oracle(conf, guide, pw, ph, sigma):
    g2 = o.nearest_resize(o.image_shift(guide, 0.5, 0.5), pw, ph)
    return o.bilateral_upsample(conf, guide, g2, sigma)


@pytest.mark.parametrize("pw,ph,W,H,sigma", [(130, 98, 260, 196, 12.0), (65, 49, 130, 98, 12.0), (1, 1, 2, 2, 12.0),
                                            (97, 31, 194, 62, 3.0), (160, 120, 320, 240, 0.75), (40, 30, 100, 60, 12.0)])
def test_upsample_layer(pw, ph, W, H, sigma):
    """Bit-exact against the oracle; the last case has a guide of another aspect than the depth (the resize steps
    differ per axis)."""
    depth, guide, conf = _upsample_inputs(pw, ph, W, H, pw * 7 + ph)
    with sgm.Context(0) as c:
        got = sgm.upsample_layer(c, depth, guide, W // 2, H // 2, sigma)
        got2, gconf = sgm.upsample_layer(c, depth, guide, W // 2, H // 2, sigma, prev_conf=conf)
    want = o.upsample_layer(depth, guide, W // 2, H // 2, sigma)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(got2.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(gconf.view(np.uint32), _upsample_conf_oracle(conf, guide, pw, ph, sigma).view(np.uint32))


def test_upsample_layer_device_and_size_rule():
    import torch

    pw, ph, W, H = 520, 390, 1040, 780
    depth, guide, conf = _upsample_inputs(pw, ph, W, H, 11)
    want = o.upsample_layer(depth, guide, W // 2, H // 2)
    want_c = _upsample_conf_oracle(conf, guide, pw, ph, 12.0)
    with sgm.Context(0) as c:
        d_prev = torch.from_numpy(depth).cuda()
        d_conf = torch.from_numpy(conf).cuda()
        d_guide = torch.from_numpy(guide).cuda()
        d_out = torch.empty((H, W), dtype=torch.float32, device="cuda")
        d_oc = torch.empty((H, W), dtype=torch.float32, device="cuda")
        for with_conf in (False, True):   # the second call reuses the held buffers
            sgm.upsample_layer_device(c, d_prev.data_ptr(), pw, ph, d_guide.data_ptr(), W, H, W // 2, H // 2,
                                      d_out.data_ptr(), d_prev_conf=d_conf.data_ptr() if with_conf else 0,
                                      d_conf=d_oc.data_ptr() if with_conf else 0)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(d_out.cpu().numpy().view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(d_oc.cpu().numpy().view(np.uint32), want_c.view(np.uint32))
        with pytest.raises(sgm.SgmError) as e:
            sgm.upsample_layer(c, depth, guide, W // 2 + 1, H // 2)
        assert e.value.code == sgm.EINVAL and "Incorrect output depth size" in str(e.value)


# ------------------------------------------------------------------ plane-sweep matching cost
from test_sweep_oracle import COLOUR, GRAY, baseline_warp, planes as sweep_planes, textured, translation_warp  # noqa: E402


def _sweep_case(seed, H0, W0, nv, n, near=500.0, far=5000.0):
    rng = np.random.default_rng(seed)
    ref = textured(H0, W0, seed)
    others, warps = [], []
    for k in range(nv):
        others.append(textured(H0 + 2 * (k % 3), W0 - 4 * (k % 4), seed * 10 + k))
        m = baseline_warp(float(rng.uniform(2000, 12000)))
        m[1], m[4] = rng.uniform(-0.02, 0.02, 2)
        m[2], m[6] = rng.uniform(-1e-4, 1e-4, 2)
        m[13] = rng.uniform(-200, 200)
        warps.append(m)
    wqs = [COLOUR if k % 2 == 0 else GRAY for k in range(nv)]
    return ref, others, warps, wqs, sweep_planes(n, near, far)


@pytest.mark.parametrize("seed,H0,W0,nv,n,scale,align,mode", [
    (1, 64, 96, 1, 24, 4, 8, "uniform"), (2, 96, 128, 3, 40, 8, 8, "ranged"), (3, 80, 100, 5, 200, 16, 8, "uniform"),
    (4, 60, 90, 2, 64, 2, 2, "ranged"), (5, 70, 70, 4, 100, 1, 8, "ranged-skip"), (6, 40, 50, 16, 16, 4, 8, "uniform"),
    (7, 50, 64, 2, 700, 8, 8, "uniform")])
def test_sweep_cost_matches_oracle(seed, H0, W0, nv, n, scale, align, mode):
    ref, others, warps, wqs, pl = _sweep_case(seed, H0, W0, nv, n)
    W, H = -(-W0 // scale), -(-H0 // scale)
    start = count = skip = None
    rng = np.random.default_rng(seed + 100)
    if mode.startswith("ranged"):
        start = rng.integers(0, n - 1, (H, W)).astype(np.uint16)
        count = rng.integers(0, np.minimum(n - start, 70) + 1).astype(np.uint16)
    if mode.endswith("skip"):
        skip = sgm.skip_mask(2, W, H)
    cc = sgm.CostContext(ref, others, warps, wqs, pl)
    with sgm.Context(0) as c:
        got = sgm.sweep_cost(c, cc, W, H, scale, align, start, count, skip)
    sc = o.SweepContext(ref, others, warps, wqs, pl)
    st = np.zeros((H, W), np.uint16) if start is None else start
    cn = np.full((H, W), n, np.uint16) if count is None else count
    want = o.layer_sweep_cost(sc, W, H, scale, align, st, cn, skip)
    np.testing.assert_array_equal(got, want)


def _scene_views(seed, H0, W0, nv, n, near=800.0, far=6000.0):
    """A reference image and cameras looking at a fronto-parallel scene at a few depths (shifts by f*B/z),
    box-filtered like DenseStereoPyramid does, with rectified-pair warps of several baselines."""
    rng = np.random.default_rng(seed)
    base = textured(H0, W0 + 64, seed)
    ref = base[:, :W0].copy()
    others, warps = [], []
    for k in range(nv):
        fb = float(rng.uniform(3000, 9000))
        shift = int(round(fb / 2500.0))
        img = base[:H0, shift:shift + W0].copy() if k % 2 == 0 else base[:H0, :W0].copy()
        others.append(img)
        m = baseline_warp(-fb if k % 2 == 0 else fb * 0.3)
        warps.append(m)
    box = lambda im: np.stack([o.box3(np.ascontiguousarray(im[:, :, c]), 255) for c in range(4)], axis=2)
    ref = box(ref)
    others = [box(im) for im in others]
    wqs = [COLOUR if k % 3 else GRAY for k in range(nv)]
    return ref, others, warps, wqs, sweep_planes(n, near, far)


@pytest.mark.parametrize("seed,H0,W0,nv,n,scale,tile", [(1, 192, 256, 3, 64, 8, 4), (2, 100, 140, 2, 200, 4, 8),
                                                        (3, 64, 96, 5, 24, 1, 16)])
def test_dense_layer_uniform(seed, H0, W0, nv, n, scale, tile):
    ref, others, warps, wqs, pl = _scene_views(seed, H0, W0, nv, n)
    cc = sgm.CostContext(ref, others, warps, wqs, pl)
    p = sgm.dense_params(W0, H0, scale, tile)
    with sgm.Context(0) as c:
        got = sgm.dense_layer(c, cc, p, want_agg=True, keep_pass1=True)
    sc = o.SweepContext(ref, others, warps, wqs, pl)
    want = o.dense_layer(sc, p.width, p.height, scale, tile, 8, gw=sgm.L16_GUIDANCE, keep_pass1=True)
    np.testing.assert_array_equal(got["agg_pass1"], want["agg_pass1"])
    np.testing.assert_array_equal(got["agg"], want["agg"])
    np.testing.assert_array_equal(got["wta"], want["wta"])
    np.testing.assert_array_equal(got["depth"].view(np.uint32), want["depth"].view(np.uint32))


@pytest.mark.parametrize("seed,H0,W0,nv,n,scale,tile,align,skipmode", [(4, 96, 128, 3, 120, 2, 8, 8, 0),
                                                                      (5, 80, 120, 2, 90, 1, 16, 8, 2),
                                                                      (6, 72, 100, 4, 60, 2, 4, 2, 2)])
def test_dense_layer_ranged(seed, H0, W0, nv, n, scale, tile, align, skipmode):
    ref, others, warps, wqs, pl = _scene_views(seed, H0, W0, nv, n)
    p = sgm.dense_params(W0, H0, scale, tile, align)
    W, H = p.width, p.height
    rng = np.random.default_rng(seed)
    count = rng.integers(0, 40, (H, W)).astype(np.uint16)
    start = rng.integers(7, n - count + 1).astype(np.uint16)         # MinLo >= 7: no padding past the segment
    skip = sgm.skip_mask(skipmode, W, H)
    cc = sgm.CostContext(ref, others, warps, wqs, pl)
    with sgm.Context(0) as c:
        got = sgm.dense_layer(c, cc, p, start, count, skip, want_agg=True)
    sc = o.SweepContext(ref, others, warps, wqs, pl)
    want = o.dense_layer(sc, W, H, scale, tile, align, start, count, skip, gw=sgm.L16_GUIDANCE)
    np.testing.assert_array_equal(got["agg"], want["agg"])
    np.testing.assert_array_equal(got["wta"], want["wta"])
    np.testing.assert_array_equal(got["depth"].view(np.uint32), want["depth"].view(np.uint32))


@pytest.mark.parametrize("seed,H0,W0,nv,n,scale,tile,align,skipmode", [(7, 96, 128, 3, 120, 2, 8, 8, 0),
                                                                      (8, 80, 120, 2, 90, 1, 15, 8, 2),
                                                                      (9, 72, 100, 4, 60, 2, 4, 2, 2)])
def test_dense_layer_spilling_padding(seed, H0, W0, nv, n, scale, tile, align, skipmode):
    """MinLo 0 and ranges ending at the top plane with padding: the line-buffer-literal path, bit for bit."""
    ref, others, warps, wqs, pl = _scene_views(seed, H0, W0, nv, n)
    p = sgm.dense_params(W0, H0, scale, tile, align)
    W, H = p.width, p.height
    rng = np.random.default_rng(seed)
    count = rng.integers(1, 40, (H, W)).astype(np.int64)
    start = rng.integers(0, n - count + 1)
    top = rng.random((H, W)) < 0.3
    start[top] = n - count[top]
    start[0, 0] = 0
    start, count = start.astype(np.uint16), count.astype(np.uint16)
    assert _spills(start, count, align) > 0
    skip = sgm.skip_mask(skipmode, W, H)
    cc = sgm.CostContext(ref, others, warps, wqs, pl)
    print(f"dense spill W={W} H={H} n={n} tile={tile} align={align}", flush=True)
    with sgm.Context(0) as c:
        got = sgm.dense_layer(c, cc, p, start, count, skip, want_agg=True)
    sc = o.SweepContext(ref, others, warps, wqs, pl)
    want = o.dense_layer(sc, W, H, scale, tile, align, start, count, skip, gw=sgm.L16_GUIDANCE)
    np.testing.assert_array_equal(got["agg"], want["agg"])
    np.testing.assert_array_equal(got["wta"], want["wta"])
    np.testing.assert_array_equal(got["depth"].view(np.uint32), want["depth"].view(np.uint32))


@pytest.mark.parametrize("seed,pw,ph,w,h,n,skipped", [(1, 16, 12, 32, 24, 64, False), (2, 65, 49, 130, 98, 300, True),
                                                       (3, 9, 7, 9, 7, 20, True), (4, 40, 30, 81, 60, 1000, True)])
def test_range_from_previous(seed, pw, ph, w, h, n, skipped):
    rng = np.random.default_rng(seed)
    pl = sweep_planes(n)
    prev = pl[rng.integers(0, n, (ph, pw))].astype(np.float32)
    jitter = rng.random((ph, pw)) < 0.5
    prev[jitter] *= rng.uniform(0.97, 1.03, int(jitter.sum())).astype(np.float32)   # between planes
    prev[0, 0] = 0.0                                                                  # rcpss(0) = inf
    pskip = sgm.skip_mask(2, pw, ph) if skipped else None
    with sgm.Context(0) as c:
        got = sgm.range_from_previous(c, pl, prev, pskip, w, h)
    want = o.range_from_previous(pl, prev, pskip, w, h)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    assert got[2:] == want[2:]


@pytest.mark.parametrize("W0,H0,seed", [(256, 192, 1), (208, 152, 2)])
def test_dense_pyramid_matches_oracle(W0, H0, seed):
    """DenseStereoPyramid.Run end to end: six levels of depth, bit for bit, and the memory mode of each level."""
    imgs, gray, warps, _ = scene.multiview(W0, H0, seed=seed)
    p8, p2 = scene.plane_list(500.0, 20000.0, 48), scene.plane_list(500.0, 20000.0, 46)
    want = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE)
    for lv in want[1:]:   # the scene keeps every range clear of the unsupported padding case
        assert lv["minlo"] >= 7
    import torch

    with sgm.Context(0) as c:
        depths, aligns = sgm.dense_pyramid(c, imgs, gray, warps, p8, p2)
        d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
        d_depths = [torch.full(d.shape, -1.0, dtype=torch.float32, device="cuda") for d in depths]
        aligns_d = sgm.dense_pyramid_device(c, [t.data_ptr() for t in d_imgs], [im.shape[:2] for im in imgs], gray, warps,
                                            p8, p2, [t.data_ptr() for t in d_depths])
        torch.cuda.synchronize()
    assert aligns == [lv["align"] for lv in want] == aligns_d
    for i, (d, dd, lv) in enumerate(zip(depths, d_depths, want)):
        np.testing.assert_array_equal(d.view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")
        np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")


@pytest.mark.parametrize("W0,H0,seed,near,far", [(256, 192, 1, 1000.0, 2200.0), (208, 152, 2, 1000.0, 2200.0),
                                                 (256, 192, 1, 1500.0, 2250.0)])
def test_dense_pyramid_spilling_padding(W0, H0, seed, near, far):
    """A plane range whose far end sits on the background: MinLo 0 on the finer levels and ranges ending at the top
    plane, so padding runs past the line-buffer segments as in real captures. Six levels, bit for bit."""
    imgs, gray, warps, _ = scene.multiview(W0, H0, seed=seed)
    p8, p2 = scene.plane_list(near, far, 48), scene.plane_list(near, far, 46)
    want = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE)
    assert sum(_spills(lv["start"], lv["count"], lv["align"]) for lv in want[1:]) > 0
    import torch

    print(f"dense pyramid spill {W0}x{H0} planes {near}..{far}", flush=True)
    with sgm.Context(0) as c:
        depths, aligns = sgm.dense_pyramid(c, imgs, gray, warps, p8, p2)
        d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
        d_depths = [torch.full(d.shape, -1.0, dtype=torch.float32, device="cuda") for d in depths]
        aligns_d = sgm.dense_pyramid_device(c, [t.data_ptr() for t in d_imgs], [im.shape[:2] for im in imgs], gray, warps,
                                            p8, p2, [t.data_ptr() for t in d_depths])
        torch.cuda.synchronize()
    assert aligns == [lv["align"] for lv in want] == aligns_d
    for i, (d, dd, lv) in enumerate(zip(depths, d_depths, want)):
        np.testing.assert_array_equal(d.view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")
        np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")


@pytest.mark.parametrize("W0,H0,seed,near,far,density", [(256, 192, 1, 500.0, 20000.0, 3.0),
                                                         (208, 152, 2, 1000.0, 2200.0, 10.0)])
def test_dense_pyramid_from_calibrations(W0, H0, seed, near, far, density):
    """DenseStereoPyramid.Run from calibrations: plane lists and warps from the library's DenseStereo.Planes and
    Mat4D.FlowMatrix equal the oracle's bit for bit, and six levels of depth computed from them equal the oracle's."""
    imgs, gray, _, _ = scene.multiview(W0, H0, seed=seed)
    cams = scene.multiview_calibs(W0, H0)
    p8, p2 = o.planes(near, far, cams, density, 8), o.planes(near, far, cams, density, 2)
    warps = [o.flow_matrix(cams[0], c) for c in cams[1:]]
    want = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE)
    print(f"dense pyramid from calibrations {W0}x{H0} planes {near}..{far} ({len(p8)}/{len(p2)})", flush=True)
    with sgm.Context(0) as c:
        depths, aligns, g8, g2, gw = sgm.dense_pyramid_run(c, imgs, gray, cams, near, far, density)
    np.testing.assert_array_equal(g8.view(np.uint32), p8.view(np.uint32))
    np.testing.assert_array_equal(g2.view(np.uint32), p2.view(np.uint32))
    np.testing.assert_array_equal(np.asarray(gw).view(np.uint32), np.asarray(warps).view(np.uint32))
    assert aligns == [lv["align"] for lv in want]
    for i, (d, lv) in enumerate(zip(depths, want)):
        np.testing.assert_array_equal(d.view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")


def test_dense_pyramid_wide_ranges():
    """304 planes over a narrow depth span: ranges up to 288 planes at the depth edges (64-plane chunks)."""
    imgs, gray, warps, _ = scene.multiview(256, 192, seed=7)
    p8, p2 = scene.plane_list(1000.0, 2200.0, 304), scene.plane_list(1000.0, 2200.0, 302)
    want = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE)
    assert max(int(lv["count"].max()) for lv in want[1:]) > 256
    print("dense pyramid wide ranges 256x192, 304 planes", flush=True)
    with sgm.Context(0) as c:
        depths, aligns = sgm.dense_pyramid(c, imgs, gray, warps, p8, p2)
    assert aligns == [lv["align"] for lv in want]
    for i, (d, lv) in enumerate(zip(depths, want)):
        np.testing.assert_array_equal(d.view(np.uint32), lv["depth"].view(np.uint32), err_msg=f"level {i}")


def test_dense_pyramid_memory_modes():
    """A budget between the last level's align-8 and align-2 estimates switches that level to align 2 with the
    align-2 planes (the smaller levels fit at align 8); a budget below both fails like the reference."""
    imgs, gray, warps, _ = scene.multiview(128, 96, seed=3)
    p8, p2 = scene.plane_list(500.0, 20000.0, 48), scene.plane_list(500.0, 20000.0, 46)
    base = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE)
    image_bytes = sum(im.shape[0] * im.shape[1] * 4 for im in imgs)
    lv = base[-1]
    est8 = o.memory_estimate(lv["W"], lv["H"], lv["start"], lv["count"], lv["skip"], 8, lv["maxhi"], image_bytes)
    budget = est8      # the last level at align 8 does not fit (>=), so it takes align 2
    want = o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE, budget=budget)
    assert [w["align"] for w in want] == [8, 8, 8, 8, 8, 2]
    p = sgm.dense_pyramid_defaults()
    p.budget = budget
    with sgm.Context(0) as c:
        depths, aligns = sgm.dense_pyramid(c, imgs, gray, warps, p8, p2, params=p)
        assert aligns == [w["align"] for w in want]
        for d, w in zip(depths, want):
            np.testing.assert_array_equal(d.view(np.uint32), w["depth"].view(np.uint32))
        p.budget = 1
        with pytest.raises(sgm.NotSupported):
            sgm.dense_pyramid(c, imgs, gray, warps, p8, p2, params=p)


def test_plane_sweep_entry_points_refuse_bad_input():
    """Argument errors come back as EINVAL before any device work; the shapes the device does not model as
    EUNSUPPORTED (the reference raises or reads past its buffers there)."""
    ref, others, warps, wqs, pl = _sweep_case(9, 40, 48, 2, 16)
    with sgm.Context(0) as c:
        cc = sgm.CostContext(ref, others, warps, wqs, pl)
        with pytest.raises(sgm.SgmError) as e:   # a range past the last plane
            sgm.sweep_cost(c, cc, 12, 10, 4, 8, np.full((10, 12), 10), np.full((10, 12), 7))
        assert e.value.code == sgm.EINVAL
        tiny = sgm.CostContext(ref, [others[0][:3, :3]], warps[:1], wqs[:1], pl)
        with pytest.raises(sgm.SgmError) as e:   # a camera image below 4 x 4
            sgm.sweep_cost(c, tiny, 12, 10, 4)
        assert e.value.code == sgm.EINVAL
        many = sgm.CostContext(ref, [others[0]] * 17, [warps[0]] * 17, [wqs[0]] * 17, pl)
        with pytest.raises(sgm.SgmError) as e:   # more than 16 cameras
            sgm.sweep_cost(c, many, 12, 10, 4)
        assert e.value.code == sgm.EINVAL
        p = sgm.dense_params(48, 40, 4, 4)
        p.width += 1
        with pytest.raises(sgm.SgmError) as e:   # a layer size that is not (ref + scale - 1) / scale
            sgm.dense_layer(c, cc, p)
        assert e.value.code == sgm.EINVAL
        big = sgm.CostContext(ref, others, warps, wqs, sweep_planes(1030))
        with pytest.raises(sgm.NotSupported):   # layer 0 above 1024 planes
            sgm.dense_layer(c, big, sgm.dense_params(48, 40, 4, 4))
        # ranges above 1024 lanes
        p = sgm.dense_params(48, 40, 4, 4)
        start = np.zeros((10, 12), np.uint16)
        count = np.full((10, 12), 3, np.uint16)
        count[5, 5] = 1030
        with pytest.raises(sgm.NotSupported):
            sgm.dense_layer(c, sgm.CostContext(ref, others, warps, wqs, sweep_planes(1040)), p, start, count)
