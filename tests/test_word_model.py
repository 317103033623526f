"""The oracle's line-buffer arithmetic pinned word for word: ``sgm_model.word_layer`` restates DenseLayer's buffers
(AllocLineBuffers / LIdx, DenseStereoSgm.cs:158-166), the stale cleanup (:254-264), the writes (:276-316) and Rd
(:318-324) from the reference's text, sharing no code with oracle/lux_oracle.c, and the two must agree on layers whose
padding spills past its segment: into the next path's segment, from path 3 into the next slot's lead word and path-0
segment, and reads past the end of the buffer (which the reference defines as 0). The same layers run on the device
in test_gpu_parity.py (SPILL_LAYERS, the same seeds), so the device's line-buffer-literal path is held to the
reference's words, not only to the oracle's."""
import numpy as np
import pytest

import sgm_model as M
from oracle import oracle as o

L16_C = np.array(0x3ef6384f, np.uint32).view(np.float32)[()]
L16_WEIGHTS = (L16_C / np.float32(6), L16_C / np.float32(16), L16_C / np.float32(16), np.float32(0))


def _layer(W, H, T, planes, cmax, align, adaptive, seed=None):
    """The layer test_gpu_parity.test_layer_ranges_spilling_padding builds for these parameters."""
    rng = np.random.default_rng(W * 1000 + H + T if seed is None else seed)
    start, count = M.spill_ranges(rng, W, H, planes, cmax, align)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raws = []
    for i in range(count.size):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    guid = rng.integers(0, 256, (H, W, 4)).astype(np.uint8) if adaptive else None
    return start, count, raws, guid


def _both(W, H, T, planes, cmax, align, adaptive, past_end=0, seed=None, xsw=False, skip=None):
    start, count, raws, guid = _layer(W, H, T, planes, cmax, align, adaptive, seed)
    kw = dict(T=T, align=align, keep_pass1=True, xsw=xsw, skip=skip)
    if adaptive:
        kw.update(p1=1, guidance=guid, gw=L16_WEIGHTS, p2_scale=500.0, adaptive=True)
        got = M.word_layer(raws, start, count, align, T, 1, 0, xsw=xsw, skip=skip, guidance=guid, gw=L16_WEIGHTS,
                           p2_scale=500.0, past_end=past_end)
    else:
        kw.update(p1=8, p2=96)
        got = M.word_layer(raws, start, count, align, T, 8, 96, xsw=xsw, skip=skip, past_end=past_end)
    want = o.layer(np.concatenate(raws), W, H, start, count, **kw)
    return got, want


def _equal(got, want):
    agg, agg1, wta, _ = got
    return (np.array_equal(agg, want[0]) and np.array_equal(agg1, want[1]) and
            np.array_equal(wta.reshape(-1), np.asarray(want[2]).reshape(-1)))


# the device's spilling layers (test_gpu_parity.SPILL_LAYERS) small enough for the pure-Python model, align 8
SPILL = [(9, 7, 4, 40, 13, 8, False), (13, 11, 3, 24, 11, 8, False), (41, 33, 7, 64, 61, 8, False)]


@pytest.mark.parametrize("W,H,T,planes,cmax,align,adaptive", SPILL)
def test_word_model_equals_oracle_on_the_device_spilling_layers(W, H, T, planes, cmax, align, adaptive):
    got, want = _both(W, H, T, planes, cmax, align, adaptive)
    assert _equal(got, want)
    ev = got[3]
    assert ev["next_segment"] > 0 and ev["foreign_read"] > 0  # spilled words are written and read back


@pytest.mark.parametrize("W,H,T,planes,cmax,adaptive,seed", [(6, 5, 2, 24, 13, False, 1), (12, 9, 4, 40, 17, True, 2),
                                                            (5, 9, 2, 16, 9, False, 3), (16, 12, 5, 30, 9, True, 4),
                                                            (7, 7, 2, 24, 5, False, 5), (11, 6, 3, 32, 30, True, 6)])
def test_word_model_equals_oracle_on_spilling_layers(W, H, T, planes, cmax, adaptive, seed):
    got, want = _both(W, H, T, planes, cmax, 8, adaptive, seed=seed)
    assert _equal(got, want)
    ev = got[3]
    assert ev["guard"] > 0 and ev["next_segment"] > 0  # padding reaches the guard word and the next segment
    assert ev["next_lead"] > 0 and ev["next_slot_path0"] > 0  # path 3 into the next slot's lead word and path 0


def test_word_model_with_xsw_and_skip_mask():
    skip = np.zeros((9, 13), np.uint8)
    skip[::2, 1::2] = 0xFF
    for xsw, sk in ((True, None), (False, skip), (True, skip)):
        got, want = _both(13, 9, 4, 32, 15, 8, False, seed=11, xsw=xsw, skip=sk)
        assert _equal(got, want), (xsw, sk is not None)


def test_past_end_reads_are_zero():
    """A layer whose up-right neighbour reads run past the end of the buffer: the last column, an odd tile row (half
    0 writes, half 1 is read: the margin slot W+1 of half 1 is the buffer's last), padding past MaxHi. The reference
    reads 0 there (Rd, :318-324); the oracle agrees with that rule and not with a 2000 guard in its place."""
    W, H, T = 3, 4, 2
    start = np.zeros((H, W), np.uint16)
    count = np.full((H, W), 8, np.uint16)
    start[1, W - 1], count[1, W - 1] = 11, 5  # cap 8 from plane 11: 19 > MaxHi 16
    start[3, W - 1], count[3, W - 1] = 11, 5
    start[0, 0], count[0, 0] = 0, 16
    rng = np.random.default_rng(7)
    caps = ((count.astype(np.int64) + 7) // 8) * 8
    raws = []
    for i in range(W * H):
        r = rng.integers(0, 49, int(caps.flat[i])).astype(np.uint8)
        r[int(count.flat[i]):] = 255
        raws.append(r)
    want = o.layer(np.concatenate(raws), W, H, start, count, T=T, align=8, p1=8, p2=96, keep_pass1=True)
    got = M.word_layer(raws, start, count, 8, T, 8, 96)
    assert got[3]["past_end"] > 0
    assert _equal(got, want)
    other = M.word_layer(raws, start, count, 8, T, 8, 96, past_end=2000)
    assert not np.array_equal(other[0], want[0])  # the rule decides results: the check has teeth


def test_past_end_rule_decides_results_on_a_device_layer():
    got, want = _both(13, 11, 3, 24, 11, 8, False)  # SPILL[1], also run on the device
    assert got[3]["past_end"] > 0 and _equal(got, want)
    other, _ = _both(13, 11, 3, 24, 11, 8, False, past_end=2000)
    assert not np.array_equal(other[0], want[0])


def test_word_model_on_uniform_and_non_spilling_layers():
    """No spill: the word model, the oracle and the label model (tests/sgm_model.label_eval) agree."""
    rng = np.random.default_rng(21)
    for W, H, T, cap in ((7, 5, 2, 16), (9, 9, 4, 24), (10, 6, 3, 8)):
        raw = rng.integers(0, 49, (H, W, cap)).astype(np.uint8)
        start = np.zeros((H, W), np.uint16)
        count = np.full((H, W), cap, np.uint16)
        got = M.word_layer(list(raw.reshape(W * H, cap)), start, count, 8, T, 10, 120)
        want = o.layer(raw.reshape(-1), W, H, start, count, T=T, align=8, p1=10, p2=120, keep_pass1=True)
        assert _equal(got, want)
        assert got[3]["next_segment"] == 0 and got[3]["past_end"] == 0
        if T % 2 == 0:
            np.testing.assert_array_equal(got[0].reshape(H, W, cap), M.label_eval(raw, T, 10, 120))


def test_align2_blocks_past_cap_are_where_the_reference_throws():
    """The reference's scalar form runs whole 8-lane blocks (:302-303): with align 2 a cap that is not a multiple of 8
    indexes the pixel's aggregate past its end, which in C# throws. The oracle and the device compute lanes [0, cap)
    there; that reading of align-2 layers is this build's, not the reference's."""
    start = np.zeros((2, 3), np.uint16)
    count = np.full((2, 3), 6, np.uint16)
    raws = [np.full(6, 3, np.uint8) for _ in range(6)]
    with pytest.raises(IndexError):
        M.word_layer(raws, start, count, 2, 2, 8, 96)
