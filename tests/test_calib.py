"""The L16 layer set-up arithmetic: the rcpss / rsqrtss models (IntelApprox.cs:21-44), Homography.Sqrt
(Homography.cs:18), DenseStereo.Planes (DenseStereo.cs:41-88) and Mat4D.FlowMatrix (ViewTransform.cs:86-133).

The oracle's restatements are pinned by the instructions themselves (every one of the 2^32 inputs on an Intel host),
by closed forms (a square root, a reciprocal, inverse-depth spacing, the rectified flow of a translated camera, the
identity of a camera mapped onto itself) and by numpy double arithmetic; the library's host functions equal the
oracle bit for bit. No reference golden vectors exist for these functions (the reference's tests do not call them).
"""
import numpy as np
import pytest

from lux_amd import scene, sgm
from oracle import oracle as o


def _intel():
    try:
        return any(l.startswith("vendor_id") and "GenuineIntel" in l for l in open("/proc/cpuinfo"))
    except OSError:
        return False


def _f(b):
    return np.uint32(b).view(np.float32)


def _b(x):
    return int(np.float32(x).view(np.uint32))


def _rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return (Rz @ Ry @ Rx).astype(np.float32)


def _random_rig(rng, n):
    cams = []
    for _ in range(n):
        f = rng.uniform(800, 4000)
        K = np.array([[f, rng.uniform(-2, 2), rng.uniform(300, 2000)], [0, f * rng.uniform(0.98, 1.02),
                      rng.uniform(200, 1500)], [0, 0, 1]], np.float32)
        R = _rot(*rng.uniform(-0.3, 0.3, 3))
        T = rng.uniform(-20, 20, 3).astype(np.float32)
        cams.append((K, T, R))
    return cams


SPECIAL = [0x0, 0x80000000, 0x7f800000, 0xff800000, 0x7fc00000, 0x7f800001, 0xffc00001, 0x00000001, 0x807fffff,
           0x00800000, 0x7f7fffff, 0xff7fffff, 0x3f800000, 0x40000000, 0x3e800000, 0x7e800000, 0x0080ffff, 0xbf800000]


# ---- the instruction models ----

@pytest.mark.skipif(not _intel(), reason="the tables are Intel's: checked against the instructions on an Intel host")
@pytest.mark.parametrize("which", ["rsqrt", "rcp"])
def test_models_match_the_instruction_on_every_input(which):
    assert o.approx_mismatches(which, 0, 1 << 32) == 0


def test_rsqrt_special_cases_and_accuracy():
    assert o.rsqrt_bits(0x00000000) == 0x7f800000 and o.rsqrt_bits(0x80000000) == 0xff800000  # +-0 -> +-inf
    assert o.rsqrt_bits(0x00000007) == 0x7f800000                                           # denormal -> inf
    assert o.rsqrt_bits(0xbf800000) == 0xffc00000                                           # negative -> QNaN
    assert o.rsqrt_bits(0x7f800000) == 0                                                    # +inf -> +0
    assert o.rsqrt_bits(0x7f800001) == 0x7fc00001                                           # NaN -> quiet NaN
    xs = np.float32(2.0) ** np.arange(-120, 120, 0.37).astype(np.float32)
    r = np.array([_f(o.rsqrt_bits(_b(x))) for x in xs], np.float64)
    rel = np.abs(r * np.sqrt(xs.astype(np.float64)) - 1.0)
    assert rel.max() <= 1.5 * 2.0 ** -12                                                    # Intel's stated bound
    # even and odd exponents take different tables: 4x a value has half its reciprocal square root, to the table
    for x in (1.1, 1.7, 3.3):
        a, b = _f(o.rsqrt_bits(_b(x))), _f(o.rsqrt_bits(_b(4 * x)))
        assert a == 2 * b


def test_library_instruction_models_equal_the_oracle():
    rng = np.random.default_rng(5)
    bits = [int(b) for b in rng.integers(0, 2 ** 32, 50000, dtype=np.uint64)] + SPECIAL
    for b in bits:
        assert sgm.rsqrtss_bits(b) == o.rsqrt_bits(b), hex(b)
        assert sgm.rcpss_bits(b) == o.rcp_bits(b), hex(b)


# ---- Homography.Sqrt ----

def test_homography_sqrt_is_a_square_root():
    assert o.homography_sqrt(0.0) == 0.0
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0, 1, 2000), rng.uniform(1, 1e4, 2000), np.float32(2.0) ** rng.uniform(-100, 100,
                         2000)]).astype(np.float32)
    got = np.array([o.homography_sqrt(x) for x in xs], np.float64)
    want = np.sqrt(xs.astype(np.float64))
    # one Newton step from a 12-bit estimate: ~2^-23 relative, a few float roundings on top
    assert (np.abs(got - want) <= 4e-7 * want + 1e-38).all()
    for x in (1.0, 4.0, 9.0, 64.0, 1e6):
        assert abs(o.homography_sqrt(x) - np.sqrt(x)) <= 2 * np.spacing(np.float32(np.sqrt(x)))
    lib = np.array([sgm.homography_sqrt(x) for x in xs], np.float32)
    np.testing.assert_array_equal(lib.view(np.uint32), got.astype(np.float32).view(np.uint32))


# ---- Mat4D.FlowMatrix ----

def _flow_f64(a, b):
    def P(c):
        K, T, R = (np.asarray(v, np.float64) for v in c)
        Kd = np.eye(4)
        Kd[:3, :3] = K
        E = np.eye(4)
        E[:3, :3] = R
        E[:3, 3] = T
        return Kd @ E
    return P(b) @ np.linalg.inv(P(a))


def test_flow_matrix_closed_forms_and_numpy():
    rng = np.random.default_rng(2)
    for trial in range(20):
        a, b = _random_rig(rng, 2)
        got = o.flow_matrix(a, b)
        want = _flow_f64(a, b)
        # column-major: element 4c + r is row r, column c
        np.testing.assert_allclose(got.reshape(4, 4).T, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())
        eye = o.flow_matrix(a, a)
        np.testing.assert_allclose(eye, np.eye(4).reshape(16), atol=1e-6)
        np.testing.assert_array_equal(sgm.flow_matrix(a, b).view(np.uint32), got.view(np.uint32))
        np.testing.assert_array_equal(sgm.flow_matrix(a, a).view(np.uint32), eye.view(np.uint32))
    # a rectified pair: camera b translated by t along x with its principal point M further right maps reference
    # pixel (x, y) at inverse depth 1/z to x + M + f t / z: m0 = m5 = m10 = m15 = 1, m8 = M, m12 = f t
    cams = scene.multiview_calibs(256, 192, fbs=(4000.0, -6000.0), margin=12)
    m = o.flow_matrix(cams[0], cams[1])
    want = np.zeros(16, np.float32)
    want[[0, 5, 10, 15]] = 1.0
    want[8], want[12] = 12.0, -4000.0
    np.testing.assert_allclose(m, want, atol=1e-9)
    np.testing.assert_allclose(o.flow_matrix(cams[0], cams[2])[12], 6000.0)


# ---- DenseStereo.Planes ----

def _centre(c):
    K, T, R = (np.asarray(v, np.float64) for v in c)
    return R.T @ T                       # the reference's (t R) row-vector product


def _n_exact(near, far, cams, density):
    B = max(np.linalg.norm(_centre(cams[0]) - _centre(c)) for c in cams)
    return (1.0 / near - 1.0 / far) * density * B * float(np.asarray(cams[0][0])[0, 0])


@pytest.mark.parametrize("near,far,density,align,seed", [(500.0, 20000.0, 3.0, 8, 0), (500.0, 20000.0, 3.0, 2, 0),
                                                         (200.0, 640000.0, 2.0, 8, 1), (1000.0, 2200.0, 10.0, 1, 2),
                                                         (300.0, 900.0, 0.9, 8, 3), (1000.0, 1001.0, 50.0, 2, 4)])
def test_planes_closed_forms(near, far, density, align, seed):
    rng = np.random.default_rng(seed)
    cams = _random_rig(rng, 4)
    p = o.planes(near, far, cams, density, align)
    n = len(p)
    x = _n_exact(near, far, cams, density)
    assert abs(x - round(x)) > 1e-3, "choose a case away from an integer plane count"
    want_n = min(int(x), 4096)
    want_n = -(-want_n // align) * align
    assert n == want_n
    assert p[-1] == np.float32(near)
    assert np.all(np.diff(p) < 0)                                   # far .. near
    exact = 1.0 / np.linspace(1.0 / far, 1.0 / near, n)
    np.testing.assert_allclose(p, exact, rtol=3e-6)                 # uniform in 1/z, rcpss + one Newton step
    lib = sgm.planes(near, far, cams, density, align)
    np.testing.assert_array_equal(lib.view(np.uint32), p.view(np.uint32))


def test_planes_accumulate_in_float_as_the_reference_steps():
    """Bit for bit against an independent float32 simulation: lanes 0..3 of 1/z start at far + l * step, advance by
    4 * step per half group, take rcpss + one Newton step in groups of eight; the rest by division from a serial sum."""
    near, far, density = 400.0, 30000.0, 7.0
    cams = scene.multiview_calibs(256, 192)
    p = o.planes(near, far, cams, density, 1)
    f32 = np.float32

    def rcp_nr(a):
        x0 = _f(o.rcp_bits(_b(a)))
        return f32(x0 + x0 * (f32(1.0) - a * x0))

    x1n, x1f = rcp_nr(f32(near)), rcp_nr(f32(far))
    n = len(p)
    step = f32(f32(x1n - x1f) / f32(n - 1))
    m = (n - 1) & ~7
    v4 = step * np.arange(4, dtype=f32) + x1f
    d4 = f32(4.0) * step
    want = np.empty(n, f32)
    for i in range(0, m, 8):
        hi = (v4 + d4).astype(f32)
        want[i:i + 4] = [rcp_nr(v) for v in v4]
        want[i + 4:i + 8] = [rcp_nr(v) for v in hi]
        v4 = (hi + d4).astype(f32)
    v = f32(f32(m) * step + x1f)
    for i in range(m, n - 1):
        want[i] = f32(1.0) / v
        v = f32(v + step)
    want[-1] = f32(near)
    assert n > 17 and (n - 1) % 8 != 0          # both the grouped part and the division tail are exercised
    np.testing.assert_array_equal(p.view(np.uint32), want.view(np.uint32))


def test_planes_baseline_is_the_largest_centre_distance_from_camera_0():
    """B = max_k |C_0 - C_k| with C = R^T T; with rotated cameras R T would give another count."""
    K = np.array([[1000, 0, 320], [0, 1000, 240], [0, 0, 1]], np.float32)
    R1, R2 = _rot(0.0, 0.6, 0.0), _rot(0.5, 0.0, 0.4)
    cams = [(K, np.array([1, 2, 3], np.float32), R1), (K, np.array([-3, 1, 0.5], np.float32), R2),
            (K, np.array([2, -2, 1], np.float32), np.eye(3, dtype=np.float32))]
    near, far, density = 500.0, 5000.0, 2.0
    n = len(o.planes(near, far, cams, density, 1))
    x = _n_exact(near, far, cams, density)
    assert n == int(x)
    wrong = max(np.linalg.norm(np.asarray(cams[0][2], np.float64) @ np.asarray(cams[0][1], np.float64)
                               - np.asarray(c[2], np.float64) @ np.asarray(c[1], np.float64)) for c in cams)
    assert int((1 / near - 1 / far) * density * wrong * 1000) != n      # the test tells R^T T from R T
    # only camera 0's focal length counts
    cams2 = [cams[0]] + [(K * np.float32(3.0), T, R) for K, T, R in cams[1:]]
    assert len(o.planes(near, far, cams2, density, 1)) == n


def test_planes_cap_and_errors():
    cams = scene.multiview_calibs(256, 192)
    assert len(o.planes(10.0, 1e6, cams, 1000.0, 8)) == 4096            # capped at 4096
    assert len(o.planes(10.0, 1e6, cams, 3e38, 8)) == 4096              # an overflowing count saturates, then the cap
    assert len(sgm.planes(10.0, 1e6, cams, 3e38, 8)) == 4096
    with pytest.raises(ValueError, match="near >= far"):
        o.planes(500.0, 500.0, cams, 2.0, 8)
    with pytest.raises(ValueError, match="non-positive"):
        o.planes(-1.0, 500.0, cams, 2.0, 8)
    with pytest.raises(ValueError, match="baseline too small"):
        o.planes(500.0, 20000.0, cams[:1], 2.0, 8)                      # one camera: B = 0
    with pytest.raises(ValueError, match="baseline too small"):
        o.planes(500.0, 20000.0, cams, float("nan"), 8)                 # a NaN count converts to 0
    with pytest.raises(ValueError, match="baseline too small"):
        o.planes(500.0, 20000.0, cams, 0.001, 8)
    for args, msg in [((500.0, 500.0, cams, 2.0, 8), "The near end has to be smaller than the fard end!"),
                      ((0.0, 500.0, cams, 2.0, 8), "Both the near and far end have to be positive!"),
                      ((500.0, 20000.0, cams[:1], 2.0, 8), "Baseline too small, check calibration.")]:
        with pytest.raises(sgm.SgmError, match=msg.replace("!", "").replace(".", r"\.")):
            sgm.planes(*args)
    with pytest.raises(sgm.SgmError):
        sgm.planes(500.0, 20000.0, cams, 2.0, 0)                       # align 0: the reference divides by zero


def test_multiview_calibs_reproduce_the_scene_warps():
    """The calibrations of the synthetic rig give, through FlowMatrix, the warps the scene's images were made with."""
    for W0, H0 in ((256, 192), (208, 152)):
        _, _, warps, _ = scene.multiview(W0, H0, seed=1)
        cams = scene.multiview_calibs(W0, H0)
        for k, w in enumerate(warps):
            np.testing.assert_allclose(o.flow_matrix(cams[0], cams[k + 1]), w, atol=1e-9)
