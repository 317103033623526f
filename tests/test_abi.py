"""CPU tests of the C ABI library: it loads, exports what include/lux_sgm.h declares, and its schedule planner
agrees with the independent label simulation of the reference schedule (no device calls)."""
import os
import re

import numpy as np
import pytest

import sgm_model as M
from lux_amd import sgm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = sgm.lib()
    text = open(os.path.join(ROOT, "include", "lux_sgm.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(lux_sgm_\w+)\s*\(", text, re.M))
    assert declared == set(sgm.EXPORTS), declared ^ set(sgm.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.lux_sgm_version() >= 100


def test_create_without_device_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(sgm.SgmError) as e:
        sgm.Context(0)
    assert e.value.code == sgm.EDEVICE
    with pytest.raises(sgm.SgmError) as e:
        sgm.Server(0, 0, slots=2)
    assert e.value.code == sgm.EDEVICE
    with pytest.raises(sgm.SgmError) as e:
        sgm.Server(0, 0, slots=0)
    assert e.value.code == sgm.EINVAL


def expected_plan(W, H, T, xsw=False):
    """Border visit lists from the full label simulation (independent of the C++ planner)."""
    lab = {}
    nvis = {}
    plus, minus = [], []
    for d, x, y, cur in M.schedule(W, H, T, xsw):
        nvis[(d, x, y)] = nvis.get((d, x, y), 0) + 1
        other = 1 - cur
        if (d > 0 and x == W - 1) or (d < 0 and x == 0):
            nbc = x - d
            codes = []
            for p, half in ((0, cur), (1, other)):
                L = lab.get((half, nbc + 1)) if 0 <= nbc < W else None
                if L is None:
                    codes.append(-1)
                elif L[0] == d:
                    assert L[1] == nbc and L[3] == 1
                    codes.append(L[2])
                else:
                    assert p == 1 and half == 0
                    codes.append(-2 - nbc)
            (plus if d > 0 else minus).append([y, cur] + codes)
        lab[(cur, x + 1)] = (d, x, y, nvis[(d, x, y)])
    return np.array(plus, np.int32).reshape(-1, 4), np.array(minus, np.int32).reshape(-1, 4)


@pytest.mark.parametrize("T", [2, 4, 8, 16])
def test_planner_matches_label_simulation(T):
    for W in sorted({1, 2, 3, T - 1, T, T + 1, 2 * T - 1, 2 * T, 2 * T + 1, 3 * T + 5} - {0}):
        for H in sorted({1, 2, 3, T - 1, T, T + 1, 2 * T + 3} - {0}):
            for xsw in (False, True):
                plus, minus, seed_row = sgm.plan_visits(W, H, T, xsw)
                eplus, eminus = expected_plan(W, H, T, xsw)
                np.testing.assert_array_equal(plus, eplus, err_msg=f"W={W} H={H} T={T} xsw={xsw}")
                np.testing.assert_array_equal(minus, eminus, err_msg=f"W={W} H={H} T={T} xsw={xsw}")
                # seed row: the last pass +1 row written to half 0
                rows = [y for y in range(H) if (y % T) % 2 == 1]
                assert seed_row == (max(rows) if rows else -1)


def test_planner_large_shapes_and_odd_tiles():
    n_plus, n_minus, seed_row = sgm.plan_info(4160, 3120, 64)
    assert n_plus >= 3120 and n_minus >= 3120 and seed_row == 3119   # 3119 % 64 = 47 is odd: half 0
    n_plus, n_minus, _ = sgm.plan_info(640, 480, 4)
    assert n_plus == 480 + 120 and n_minus == 480 + 120   # 120 double visits per pass (T=4)
    n_plus, _, _ = sgm.plan_info(640, 480, 4, xsw=True)
    assert n_plus == 480                                  # LUX_SGM_XSW=1 removes pass +1's double visits
    with pytest.raises(sgm.NotSupported):
        sgm.plan_info(64, 48, 5)


def test_mt19937_known_answers():
    """std::mt19937's defining values (C++ [rand.predef]: the 10000th output of the default seed 5489 is
    4123659995) and numpy's legacy MT19937 stream (same seeding) at several positions."""
    assert sgm.mt19937_draw(5489, 1) == 3499211612
    assert sgm.mt19937_draw(5489, 10000) == 4123659995
    stream = np.random.RandomState(7).randint(0, 2 ** 32, size=1300, dtype=np.uint32)
    for n in (1, 2, 623, 624, 625, 1248, 1300):
        assert sgm.mt19937_draw(7, n) == int(stream[n - 1])


def _skip_mask_model(w, h):
    """DenseLayer.SkipMask mode 2 written from its description with numpy's MT19937 (legacy seeding)."""
    m = np.full((h, w), 0xFF, np.uint8)
    nx = max(1, w // 64 + (1 if 64 < 2 * (w % 64) else 0))
    ny = max(1, h // 64 + (1 if 64 < 2 * (h % 64) else 0))
    for j in range(ny):
        for i in range(nx):
            x0, y0 = 64 * i, 64 * j
            x1 = min(w, x0 + 64 * (2 if i == nx - 1 else 1))
            y1 = min(h, y0 + 64 * (2 if j == ny - 1 else 1))
            draws = iter(np.random.RandomState(5489).randint(0, 2 ** 32, size=2 * 64 * 64, dtype=np.uint32))
            for y in range(y0, y1, 2):
                for x in range(x0, x1, 2):
                    dy, dx = int(next(draws)) & 1, int(next(draws)) & 1
                    if x + dx < w and y + dy < h:
                        m[y + dy, x + dx] = 0
    return m


@pytest.mark.parametrize("w,h", [(65, 49), (130, 98), (260, 195), (1, 1), (64, 64), (97, 31), (200, 150)])
def test_skip_mask_mode2(w, h):
    got = sgm.skip_mask(2, w, h)
    np.testing.assert_array_equal(got, _skip_mask_model(w, h))
    assert (sgm.skip_mask(0, w, h) == 0).all()
    with pytest.raises(sgm.NotSupported):
        sgm.skip_mask(1, w, h)


def test_rect_bytes_rejects_out_of_range_plane_counts():
    """Plane counts past the uint16 range are refused before the lane rounding, which would overflow int."""
    assert sgm.rect_bytes(640, 480, 64, bits=8) > 640 * 480 * 64 * 3
    for D, align in ((2 ** 31 - 1, 8), (65536, 8), (64, 2 ** 31 - 1)):
        with pytest.raises(sgm.NotSupported):
            sgm.rect_bytes(64, 48, D, align=align)
    with pytest.raises(sgm.NotSupported):
        sgm.rect_bytes(64, 48, 1025)


def _chain_len(c, W, H, DX):
    """Length of diagonal chain c (chain_start in kernels_impl.h): c < W starts on the first row at x = c, else on the
    first (DX > 0) or last (DX < 0) column at row c - W + 1."""
    if c < W:
        x0, y0 = c, 0
    else:
        x0, y0 = (0 if DX > 0 else W - 1), c - W + 1
    nx = W - x0 if DX > 0 else x0 + 1
    return min(nx, H - y0)


def _diag_order(w, W, H, DX):
    """diag_order<DX> of kernels_impl.h, restated: the chain wavefront w runs."""
    if w >= W + H - 1:
        return w
    if W >= H:
        F = W - H + 1
        if w < F:
            c = w
        else:
            t = w - F
            j = t // 2 + 1
            c = W - 1 + j if t & 1 else W - H + j
    else:
        F = H - W + 1
        if w < F:
            c = 0 if w == 0 else W - 1 + w
        else:
            t = w - F
            j = t // 2 + 1
            c = W - 1 + (H - W) + j if t & 1 else j
    return W - 1 - c if (DX < 0 and c < W) else c


def test_diag_order_is_a_longest_first_permutation():
    """Every chain runs once, the full-length ones first and lengths never grow, for all shapes up to 40 x 40 (wider,
    square and taller than wide) and both horizontal directions."""
    for W in range(1, 41):
        for H in range(1, 41):
            n = W + H - 1
            for DX in (1, -1):
                order = [_diag_order(w, W, H, DX) for w in range(n)]
                assert sorted(order) == list(range(n)), (W, H, DX)
                lens = [_chain_len(c, W, H, DX) for c in order]
                assert all(a >= b for a, b in zip(lens, lens[1:])), (W, H, DX, lens)
                assert lens[0] == min(W, H) and sum(lens) == W * H
                assert _diag_order(n + 3, W, H, DX) == n + 3
