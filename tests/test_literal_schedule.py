"""The line-buffer-literal schedule of lux_amd/csrc/literal.hip, checked on the CPU through its restatement in
sgm_model (literal_rows / literal_need): the rows' visit sequences are exactly the reference's visits
(DenseStereoSgm.cs:206-222) in its order, and every order the waits allow, with as few resident wavefronts as the
host requires, gives the reference's aggregates on layers whose padding spills past the line-buffer segments."""
import numpy as np
import pytest

import sgm_model as M
from oracle import oracle as o


@pytest.mark.parametrize("W,H,T,xsw", [(1, 1, 2, False), (5, 3, 2, False), (7, 9, 4, False), (7, 9, 4, True),
                                       (9, 13, 3, False), (13, 10, 5, True), (16, 9, 8, False), (17, 17, 8, True),
                                       (33, 12, 16, False), (20, 7, 64, False), (11, 19, 7, False)])
def test_rows_are_the_reference_visits(W, H, T, xsw):
    for d in (1, -1):
        ref = [(x, y, cur) for dd, x, y, cur in M.schedule(W, H, T, xsw) if dd == d]
        rows = M.literal_rows(W, H, T, xsw, d)
        assert sum(len(s) for s in rows) == len(ref)
        for row, seq in enumerate(rows):
            y = row if d > 0 else H - 1 - row
            assert [(x, cur) for x, yy, cur, _ in seq] == [(x, cur) for x, yy, cur in ref if yy == y]
            keys = [k for *_, k in seq]
            assert keys == sorted(keys)
        # sorting every visit by (key, position in its row) gives the reference's order
        tagged = sorted((k, pos, (x, y, cur)) for seq in rows for pos, (x, y, cur, k) in enumerate(seq))
        assert [v for *_, v in tagged] == ref


def _layer_inputs(rng, W, H, planes, cmax, align):
    start, count = M.spill_ranges(rng, W, H, planes, cmax, align)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    minlo = int(start.min())
    maxhi = int(-(-int((start.astype(np.int64) + count).max()) // align) * align)
    spill = int((start.astype(np.int64) - minlo + caps > maxhi).sum())
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    return start, count, np.concatenate(raws), spill


def test_visit_order_entry_in_the_reference_order():
    # lo_layer_compute_visits given the reference's own order is lo_layer_compute
    W, H, T, planes, cmax, align = 11, 9, 4, 32, 13, 8
    rng = np.random.default_rng(3)
    start, count, raw, _ = _layer_inputs(rng, W, H, planes, cmax, align)
    kw = dict(T=T, align=align, p1=8, p2=96, keep_pass1=True)
    visits = [(x, y, cur) for d, x, y, cur in M.schedule(W, H, T)]
    n_plus = sum(1 for d, *_ in M.schedule(W, H, T) if d > 0)
    a = o.layer(raw, W, H, start, count, **kw)
    b = o.layer(raw, W, H, start, count, visits=np.array(visits, np.int32), n_plus=n_plus, **kw)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


SPILL_LAYERS = [(9, 7, 4, False, 40, 13, 8), (13, 11, 4, True, 24, 11, 8), (16, 12, 3, False, 30, 9, 2),
                (12, 17, 8, False, 64, 21, 8), (21, 9, 2, False, 16, 7, 8), (6, 14, 5, True, 48, 30, 8),
                (11, 16, 4, False, 63, 21, 8), (14, 19, 2, False, 55, 39, 8), (17, 13, 6, True, 45, 17, 8)]


@pytest.mark.parametrize("W,H,T,xsw,planes,cmax,align", SPILL_LAYERS)
def test_allowed_orders_give_the_reference_results(W, H, T, xsw, planes, cmax, align):
    rng = np.random.default_rng(W * 100 + H + T)
    start, count, raw, spill = _layer_inputs(rng, W, H, planes, cmax, align)
    assert spill > 0  # the layer exercises the words past the segments
    kw = dict(T=T, align=align, p1=8, p2=96, xsw=xsw, keep_pass1=True)
    agg, agg1, wta = o.layer(raw, W, H, start, count, **kw)
    flags = M.literal_flags(start, count, align)
    for seed, resident in ((1, M.literal_depth(W, T)), (2, M.literal_depth(W, T)), (3, 10 ** 6), (4, 10 ** 6)):
        visits, n_plus = M.literal_interleave(W, H, T, xsw, np.random.default_rng(seed), resident, flags)
        got, got1, gw = o.layer(raw, W, H, start, count, visits=visits, n_plus=n_plus, **kw)
        np.testing.assert_array_equal(got1, agg1)
        np.testing.assert_array_equal(got, agg)
        np.testing.assert_array_equal(gw, wta)


def test_allowed_orders_adaptive_p2():
    W, H, T, planes, cmax, align = 14, 10, 4, 40, 15, 8
    rng = np.random.default_rng(5)
    start, count, raw, spill = _layer_inputs(rng, W, H, planes, cmax, align)
    assert spill > 0
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
    kw = dict(T=T, align=align, p1=1, guidance=guid, gw=(0.3, 0.02, 0.05, 0.0), p2_scale=40.0, adaptive=True,
              keep_pass1=True)
    agg, agg1, wta = o.layer(raw, W, H, start, count, **kw)
    visits, n_plus = M.literal_interleave(W, H, T, False, np.random.default_rng(9), M.literal_depth(W, T),
                                          M.literal_flags(start, count, align))
    got, got1, gw = o.layer(raw, W, H, start, count, visits=visits, n_plus=n_plus, **kw)
    np.testing.assert_array_equal(got1, agg1)
    np.testing.assert_array_equal(got, agg)
    np.testing.assert_array_equal(gw, wta)


def test_a_shorter_reach_breaks_the_order(monkeypatch):
    # the check has teeth: waits that reach one column less than the kernel's let orders through that do not give
    # the reference's results on a spilling layer
    W, H, T, planes, cmax, align = 13, 11, 4, 24, 11, 8
    rng = np.random.default_rng(W * 100 + H + T)
    start, count, raw, _ = _layer_inputs(rng, W, H, planes, cmax, align)
    kw = dict(T=T, align=align, p1=8, p2=96, keep_pass1=True)
    agg, _, _ = o.layer(raw, W, H, start, count, **kw)
    monkeypatch.setattr(M, "LIT_REACH", 1)
    differs = 0
    for seed in range(6):
        visits, n_plus = M.literal_interleave(W, H, T, False, np.random.default_rng(seed), 10 ** 6)
        got, _, _ = o.layer(raw, W, H, start, count, visits=visits, n_plus=n_plus, **kw)
        differs += int(not np.array_equal(got, agg))
    assert differs > 0


@pytest.mark.parametrize("drop", ["pe", "sp"])
def test_the_per_visit_reach_has_teeth(drop):
    # each of the two conditions that lengthen a visit's reach to two columns is needed: without it, orders the
    # waits would allow give other results on spilling layers
    cases = {"pe": (11, 16, 4, 63, 21, 8), "sp": (10, 20, 4, 48, 37, 8)}
    W, H, T, planes, cmax, align = cases[drop]
    differs = 0
    for trial in range(4):
        rng = np.random.default_rng(trial)
        start, count, raw, _ = _layer_inputs(rng, W, H, planes, cmax, align)
        kw = dict(T=T, align=align, p1=8, p2=96, keep_pass1=True)
        agg, _, _ = o.layer(raw, W, H, start, count, **kw)
        pe, sp = M.literal_flags(start, count, align)
        flags = (np.zeros_like(pe), sp) if drop == "pe" else (pe, np.zeros_like(sp))
        for seed in range(6):
            visits, n_plus = M.literal_interleave(W, H, T, False, np.random.default_rng(seed), 10 ** 6, flags)
            got, _, _ = o.layer(raw, W, H, start, count, visits=visits, n_plus=n_plus, **kw)
            differs += int(not np.array_equal(got, agg))
    assert differs > 0
