"""Host-side mirror of Lux.Engine's dense-stereo layer interface over liblux_sgm.so (ctypes).

``DenseLayer`` keeps the reference's public surface (DenseStereoSgm.cs:13-26, 172-174): ``Agg``, ``Raw``,
``Start``, ``Count``, ``Cap``, ``Depth``-as-plane-index, ``AggPass1``, ``Threads``, ``KeepPass1`` and
``compute()``; the work runs on the MI355X through the C ABI (include/lux_sgm.h). ``rectified`` is the
rectified-pair front end (census / AD cost -> SGM -> WTA + subpixel + L/R). There is no CPU path here:
without the built library every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblux_sgm.so")

OK, EINVAL, EUNSUPPORTED, ENOMEM, EDEVICE, ECANCELED = 0, -1, -2, -3, -4, -5
COST_CENSUS, COST_AD = 0, 1

EXPORTS = ("lux_sgm_version", "lux_sgm_create", "lux_sgm_destroy", "lux_sgm_last_error", "lux_sgm_rect_bytes",
           "lux_sgm_rectified", "lux_sgm_rectified_device",
           "lux_sgm_rectified_device_noalloc", "lux_sgm_compute_layer", "lux_sgm_plan_info",
           "lux_sgm_plan_visits", "lux_sgm_set_profiling", "lux_sgm_stage_times", "lux_sgm_set_option",
           "lux_sgm_compute_layer_ranges", "lux_sgm_compute_layer_ex", "lux_sgm_skip_mask", "lux_sgm_mt19937_draw",
           "lux_sgm_box3", "lux_sgm_box3_device", "lux_sgm_upsample_layer", "lux_sgm_upsample_layer_device",
           "lux_sgm_sweep_cost", "lux_sgm_range_from_previous", "lux_sgm_dense_layer", "lux_sgm_dense_pyramid",
           "lux_sgm_dense_pyramid_defaults", "lux_sgm_dense_pyramid_device",
           "lux_sgm_pyramid_bytes", "lux_sgm_pyramid", "lux_sgm_pyramid_device",
           "lux_sgm_batch", "lux_sgm_batch_tune", "lux_sgm_batch_set_schedule", "lux_sgm_server_create", "lux_sgm_server_create_ex",
           "lux_sgm_server_destroy",
           "lux_sgm_server_last_error", "lux_sgm_server_submit", "lux_sgm_server_submit_pyramid", "lux_sgm_server_poll",
           "lux_sgm_server_wait",
           "lux_sgm_server_cancel", "lux_sgm_server_info_get", "lux_sgm_planes", "lux_sgm_flow_matrix",
           "lux_sgm_homography_sqrt", "lux_sgm_rcpss_bits", "lux_sgm_rsqrtss_bits", "lux_sgm_calib_last_error")
MEM_HOST, MEM_DEVICE = 0, 1
REQ_PENDING, REQ_RUNNING, REQ_DONE = 0, 1, 2
SERVER_DEVICE_LOOP = 1
OPT_FUSED_BANDS = 1
OPT_LITERAL = 2
STAGES = ("census", "rows+", "vd+1", "vd+2", "vd+3", "replay+", "vd-1", "vd-2", "vd-3", "rows-", "replay-", "finish")


class SgmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lux_sgm error {code}: {msg}")
        self.code = code


class NotSupported(SgmError):
    """The reference's NotSupportedException (DenseStereoSgm.cs:69, :343)."""


class RectParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("width", "height", "stride", "bits", "cost", "census_radius", "ad_cap",
                                              "disparities", "align", "p1", "p2", "tile", "xs_clamp_w", "lr_tolerance")]


class Pair(ctypes.Structure):
    """lux_sgm_pair: one pair of a batch, every pointer device memory (0 = none)."""
    _fields_ = [("params", RectParams), ("left", ctypes.c_void_p), ("right", ctypes.c_void_p),
                ("disp", ctypes.c_void_p), ("sub", ctypes.c_void_p), ("valid", ctypes.c_void_p)]


class Schedule(ctypes.Structure):
    """lux_sgm_schedule: graph branches (executor 0) or the persistent executor (1: one launch, 7: per pass group)."""
    _fields_ = [("bands", ctypes.c_int32), ("branches", ctypes.c_int32), ("ms", ctypes.c_float),
                ("executor", ctypes.c_int32), ("order", ctypes.c_int32)]


class RequestStats(ctypes.Structure):
    _fields_ = [("queued_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("slot", ctypes.c_int32), ("status", ctypes.c_int32)]


class ServerInfo(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("pending", ctypes.c_int32), ("running", ctypes.c_int32), ("slots", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("budget", ctypes.c_uint64), ("bytes_held", ctypes.c_uint64),
                ("bytes_peak", ctypes.c_uint64)]


class PyramidParams(ctypes.Structure):
    _fields_ = [("base", RectParams), ("levels", ctypes.c_int32), ("radius", ctypes.c_int32 * 3)]


class Guidance(ctypes.Structure):
    """lux_sgm_guidance: adaptive P2 from an RGBA8 guidance image (DenseStereoSgm.cs:267-274)."""
    _fields_ = [("guidance", ctypes.c_void_p), ("weights", ctypes.c_float * 4), ("p2_scale", ctypes.c_float)]


class LayerParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("width", "height", "planes", "align", "p1", "p2", "tile", "xs_clamp_w")]


_LIB = None


def lib():
    """Load liblux_sgm.so; raises if it was not built (no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or `make -C lux_amd/csrc`")
        # PyTorch-ROCm carries its own HIP runtime whose file is named differently from the system one: load it
        # first so that the library binds to it, and one runtime (not two competing for the device) serves the
        # process when tensors and the library share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64
        L.lux_sgm_version.restype = ctypes.c_int
        L.lux_sgm_create.argtypes = [ctypes.c_int, u64, ctypes.POINTER(vp)]
        L.lux_sgm_destroy.argtypes = [vp]
        L.lux_sgm_last_error.argtypes = [vp]
        L.lux_sgm_last_error.restype = ctypes.c_char_p
        L.lux_sgm_rect_bytes.argtypes = [vp, ctypes.POINTER(u64)]
        L.lux_sgm_rectified.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_rectified_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_rectified_device_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_compute_layer.argtypes = [vp, vp, vp, vp, vp, vp]
        L.lux_sgm_compute_layer_ranges.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_compute_layer_ex.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_skip_mask.argtypes = [i32, i32, i32, vp]
        L.lux_sgm_box3.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp]
        L.lux_sgm_sweep_cost.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]
        L.lux_sgm_range_from_previous.argtypes = [vp, vp, i32, vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp]
        L.lux_sgm_dense_layer.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_dense_pyramid.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, vp, i32, vp, vp]
        L.lux_sgm_dense_pyramid_device.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, vp, i32, vp, vp]
        L.lux_sgm_dense_pyramid_defaults.argtypes = [vp]
        L.lux_sgm_dense_pyramid_defaults.restype = None
        L.lux_sgm_upsample_layer.argtypes = [vp, vp, vp, i32, i32, vp, i32, i32, i32, i32, ctypes.c_float, vp, vp]
        L.lux_sgm_upsample_layer_device.argtypes = [vp, vp, vp, i32, i32, vp, i32, i32, i32, i32, ctypes.c_float, vp, vp,
                                                    vp]
        L.lux_sgm_box3_device.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, vp]
        L.lux_sgm_planes.argtypes = [ctypes.c_float, ctypes.c_float, vp, i32, ctypes.c_float, i32, vp, i32,
                                     ctypes.POINTER(i32)]
        L.lux_sgm_flow_matrix.argtypes = [vp, vp, vp]
        L.lux_sgm_homography_sqrt.argtypes = [ctypes.c_float]
        L.lux_sgm_homography_sqrt.restype = ctypes.c_float
        L.lux_sgm_rcpss_bits.argtypes = [ctypes.c_uint32]
        L.lux_sgm_rcpss_bits.restype = ctypes.c_uint32
        L.lux_sgm_rsqrtss_bits.argtypes = [ctypes.c_uint32]
        L.lux_sgm_rsqrtss_bits.restype = ctypes.c_uint32
        L.lux_sgm_calib_last_error.argtypes = []
        L.lux_sgm_calib_last_error.restype = ctypes.c_char_p
        L.lux_sgm_mt19937_draw.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.lux_sgm_mt19937_draw.restype = ctypes.c_uint32
        L.lux_sgm_plan_info.argtypes = [i32, i32, i32, i32, vp, vp, vp]
        L.lux_sgm_plan_visits.argtypes = [i32, i32, i32, i32, vp, i32, vp, i32]
        L.lux_sgm_set_profiling.argtypes = [vp, ctypes.c_int]
        L.lux_sgm_stage_times.argtypes = [vp, vp, i32]
        L.lux_sgm_set_option.argtypes = [vp, i32, i32]
        L.lux_sgm_pyramid_bytes.argtypes = [vp, ctypes.POINTER(u64)]
        L.lux_sgm_pyramid.argtypes = [vp, vp, vp, vp, vp, vp]
        L.lux_sgm_pyramid_device.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.lux_sgm_batch.argtypes = [vp, i32, vp, vp]
        L.lux_sgm_batch_tune.argtypes = [vp, i32, vp, vp, i32, vp, vp, i32]
        L.lux_sgm_batch_set_schedule.argtypes = [vp, i32, vp, vp]
        L.lux_sgm_server_create.argtypes = [ctypes.c_int, u64, i32, ctypes.POINTER(vp)]
        L.lux_sgm_server_create_ex.argtypes = [ctypes.c_int, u64, i32, ctypes.c_uint32, ctypes.POINTER(vp)]
        L.lux_sgm_server_destroy.argtypes = [vp]
        L.lux_sgm_server_last_error.argtypes = [vp]
        L.lux_sgm_server_last_error.restype = ctypes.c_char_p
        L.lux_sgm_server_submit.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, ctypes.POINTER(u64)]
        L.lux_sgm_server_submit_pyramid.argtypes = [vp, vp, i32, vp, vp, vp, vp, ctypes.POINTER(u64)]
        L.lux_sgm_server_poll.argtypes = [vp, u64, ctypes.POINTER(i32)]
        L.lux_sgm_server_wait.argtypes = [vp, u64, i32, vp]
        L.lux_sgm_server_cancel.argtypes = [vp, u64]
        L.lux_sgm_server_info_get.argtypes = [vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().lux_sgm_last_error(ctx).decode("utf-8", "replace")
        raise (NotSupported if rc == EUNSUPPORTED else SgmError)(rc, msg)


def plan_info(W, H, T, xsw=False):
    a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib().lux_sgm_plan_info(W, H, T, int(bool(xsw)), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value


def plan_visits(W, H, T, xsw=False):
    n_plus, n_minus, seed_row = plan_info(W, H, T, xsw)
    plus = np.zeros((max(n_plus, 1), 4), np.int32)
    minus = np.zeros((max(n_minus, 1), 4), np.int32)
    _check(lib().lux_sgm_plan_visits(W, H, T, int(bool(xsw)), _p(plus), n_plus, _p(minus), n_minus))
    return plus[:n_plus], minus[:n_minus], seed_row


class Context:
    """One device context (lux_sgm_create). hbm_budget in bytes, 0 = unbounded."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        _check(lib().lux_sgm_create(device, hbm_budget, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().lux_sgm_destroy(self.handle)
            self.handle = None

    def set_option(self, option, value):
        _check(lib().lux_sgm_set_option(self.handle, option, value), self.handle)

    def set_profiling(self, on=True):
        _check(lib().lux_sgm_set_profiling(self.handle, int(bool(on))), self.handle)

    def stage_times(self):
        """Durations (ms) of the stages of the last profiled request, keyed by STAGES."""
        ms = np.zeros(16, np.float32)
        n = lib().lux_sgm_stage_times(self.handle, _p(ms), 16)
        if n < 0:
            _check(n, self.handle)
        return dict(zip(STAGES, ms[:n].tolist()))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rect_params(W, H, D, *, bits=8, cost="census", radius=3, ad_cap=28, align=8, p1=8, p2=96, tile=64, xsw=False,
                tol=1, stride=None):
    p = RectParams()
    p.width, p.height, p.stride, p.bits = W, H, stride or W, bits
    p.cost = COST_CENSUS if cost == "census" else COST_AD
    p.census_radius, p.ad_cap, p.disparities, p.align = radius, ad_cap, D, align
    p.p1, p.p2, p.tile, p.xs_clamp_w, p.lr_tolerance = p1, p2, tile, int(bool(xsw)), tol
    return p


def rect_bytes(W, H, D, **kw):
    out = ctypes.c_uint64()
    _check(lib().lux_sgm_rect_bytes(ctypes.byref(rect_params(W, H, D, **kw)), ctypes.byref(out)))
    return out.value


def rectified(ctx, left, right, D, *, want_agg=False, **kw):
    """Rectified pair from host arrays. Returns dict(disp, sub, valid[, agg])."""
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    H, W = left.shape
    bits = kw.pop("bits", 8 if left.dtype == np.uint8 else 10)
    if (left.dtype == np.uint8) != (bits == 8):
        raise ValueError("bits does not match the image dtype")
    p = rect_params(W, H, D, bits=bits, **kw)
    cap = ((D + p.align - 1) // p.align) * p.align
    disp = np.empty((H, W), np.uint16)
    sub = np.empty((H, W), np.float32)
    valid = np.empty((H, W), np.uint8)
    agg = np.empty((H, W, cap), np.uint16) if want_agg else None
    _check(lib().lux_sgm_rectified(ctx.handle, ctypes.byref(p), _p(left), _p(right), _p(disp), _p(sub), _p(valid),
                                   _p(agg)), ctx.handle)
    out = dict(disp=disp, sub=sub, valid=valid)
    if want_agg:
        out["agg"] = agg
    return out


def rectified_device(ctx, p, d_left, d_right, d_disp, d_sub, d_valid, d_agg=0, stream=0):
    """Device-pointer entry point: pointers are ints (e.g. torch tensor data_ptr()); enqueue only."""
    _check(lib().lux_sgm_rectified_device(ctx.handle, ctypes.byref(p), ctypes.c_void_p(d_left),
                                          ctypes.c_void_p(d_right), ctypes.c_void_p(d_disp), ctypes.c_void_p(d_sub),
                                          ctypes.c_void_p(d_valid), ctypes.c_void_p(d_agg or None),
                                          ctypes.c_void_p(stream or None)), ctx.handle)


def rectified_device_noalloc(ctx, p, d_left, d_right, d_disp, d_sub, d_valid, d_agg=0, stream=0):
    """lux_sgm_rectified_device_noalloc: the same request with its buffers in the library's static scratch."""
    _check(lib().lux_sgm_rectified_device_noalloc(ctx.handle, ctypes.byref(p), ctypes.c_void_p(d_left),
                                                  ctypes.c_void_p(d_right), ctypes.c_void_p(d_disp),
                                                  ctypes.c_void_p(d_sub), ctypes.c_void_p(d_valid),
                                                  ctypes.c_void_p(d_agg or None), ctypes.c_void_p(stream or None)),
           ctx.handle)


class DenseLayer:
    """Mirror of DenseLayer (DenseStereoSgm.cs:13-26).

    Uniform range (the layer-0 form of Init): set ``Raw`` (uint8 [H, W, Cap]) and call ``compute()``;
    ``Agg`` (uint16 [H, W, Cap]), ``AggPass1`` (with ``KeepPass1``) and ``Wta`` (plane index start+arg, the
    argument of ``Depth = Planes[...]``, :225-231) are filled from the device.
    Per-pixel ranges (as RangeFromPrevious leaves them, :91-140): call ``set_ranges(start, count)`` first;
    ``Raw`` is then the jagged volume (``Cap[i]`` bytes per pixel, concatenated in raster order) and ``Agg`` /
    ``AggPass1`` come back the same way.
    """

    def __init__(self, ctx, W, H, planes, *, align=8, p1=1, p2=0, tile=4, xsw=False):
        self.ctx, self.W, self.H, self.Align, self.Tile, self.Xsw = ctx, W, H, align, tile, xsw
        self.Planes = planes
        self.P1, self.P2 = p1, p2
        cap = ((planes + align - 1) // align) * align
        self.Start = np.zeros(W * H, np.uint16)
        self.Count = np.full(W * H, planes, np.uint16)
        self.Cap = np.full(W * H, cap, np.uint16)
        self.MinLo, self.MaxHi = 0, cap
        self.Raw = None
        self.Agg = None
        self.AggPass1 = None
        self.Wta = None
        self.Threads = 1
        self.KeepPass1 = False

    def set_guidance(self, guidance, weights, p2_scale):
        """Adaptive P2: Guidance (uint8 [H, W, 4]), P.Guidance weights and P.P2Scale (StereoParams,
        DenseStereo.cs:15-31); every path step then takes P2 = (int)(P1 * P2Scale * fastexp2(-S))."""
        self.Guidance = np.ascontiguousarray(guidance, np.uint8).reshape(self.H, self.W, 4)
        self.GuidanceWeights = tuple(float(w) for w in weights)
        self.P2Scale = float(p2_scale)
        return self

    def _guidance_arg(self):
        g = getattr(self, "Guidance", None)
        if g is None:
            return None
        gd = Guidance()
        gd.guidance = g.ctypes.data
        for k in range(4):
            gd.weights[k] = self.GuidanceWeights[k]
        gd.p2_scale = self.P2Scale
        return ctypes.byref(gd)

    def _skip_arg(self):
        sk = getattr(self, "Skip", None)
        if sk is None:
            return None
        self.Skip = np.ascontiguousarray(sk, np.uint8).reshape(-1)
        return _p(self.Skip)

    def set_ranges(self, start, count):
        self.Start = np.ascontiguousarray(start, np.uint16).reshape(-1)
        self.Count = np.ascontiguousarray(count, np.uint16).reshape(-1)
        self.Cap = (((self.Count.astype(np.int64) + self.Align - 1) // self.Align) * self.Align).astype(np.uint16)
        self.MinLo = int(self.Start.min())
        self.MaxHi = int((self.Start.astype(np.int64) + self.Count).max() + self.Align - 1) // self.Align * self.Align
        self._ranged = True
        return self

    def compute(self):
        if getattr(self, "_ranged", False):
            return self._compute_ranges()
        cap = int(self.Cap[0])
        raw = np.ascontiguousarray(self.Raw, np.uint8).reshape(self.H, self.W, cap)
        p = LayerParams()
        p.width, p.height, p.planes, p.align = self.W, self.H, int(self.Count[0]), self.Align
        p.p1, p.p2, p.tile, p.xs_clamp_w = self.P1, self.P2, self.Tile, int(bool(self.Xsw))
        self.Agg = np.empty((self.H, self.W, cap), np.uint16)
        self.AggPass1 = np.empty_like(self.Agg) if self.KeepPass1 else None
        self.Wta = np.empty((self.H, self.W), np.uint16)
        _check(lib().lux_sgm_compute_layer_ex(self.ctx.handle, ctypes.byref(p), None, None, self._skip_arg(),
                                              self._guidance_arg(), _p(raw), _p(self.Agg), _p(self.AggPass1),
                                              _p(self.Wta)), self.ctx.handle)
        return self

    def _compute_ranges(self):
        total = int(self.Cap.astype(np.int64).sum())
        raw = np.ascontiguousarray(self.Raw, np.uint8).reshape(-1)
        if raw.size != total:
            raise ValueError(f"Raw holds {raw.size} bytes, the ranges need {total}")
        p = LayerParams()
        p.width, p.height, p.planes, p.align = self.W, self.H, self.Planes, self.Align
        p.p1, p.p2, p.tile, p.xs_clamp_w = self.P1, self.P2, self.Tile, int(bool(self.Xsw))
        self.Agg = np.empty(max(total, 1), np.uint16)[:total]
        self.AggPass1 = np.empty_like(self.Agg) if self.KeepPass1 else None
        self.Wta = np.empty((self.H, self.W), np.uint16)
        _check(lib().lux_sgm_compute_layer_ex(self.ctx.handle, ctypes.byref(p), _p(self.Start), _p(self.Count),
                                              self._skip_arg(), self._guidance_arg(), _p(raw), _p(self.Agg),
                                              _p(self.AggPass1), _p(self.Wta)), self.ctx.handle)
        return self


def batch(ctx, pairs, stream=0):
    """Enqueue a list of Pair as one graph launch (lux_sgm_batch)."""
    arr = (Pair * len(pairs))(*pairs)
    _check(lib().lux_sgm_batch(ctx.handle, len(pairs), arr, ctypes.c_void_p(stream or None)), ctx.handle)


def batch_set_schedule(ctx, pairs, *, executor=0, order=0, branches=4, bands=0):
    """Use this schedule for batches of these pairs' shapes (lux_sgm_batch_set_schedule)."""
    arr = (Pair * len(pairs))(*pairs)
    sch = Schedule(bands, branches, 0.0, executor, order)
    _check(lib().lux_sgm_batch_set_schedule(ctx.handle, len(pairs), arr, ctypes.byref(sch)), ctx.handle)


def batch_tune(ctx, pairs, stream=0, reps=2):
    """Time every candidate schedule on this batch; returns (best, [tried...]) as dicts."""
    arr = (Pair * len(pairs))(*pairs)
    best = Schedule()
    tried = (Schedule * 32)()
    n = lib().lux_sgm_batch_tune(ctx.handle, len(pairs), arr, ctypes.c_void_p(stream or None), reps,
                                 ctypes.byref(best), tried, 32)
    if n < 0:
        _check(n, ctx.handle)
    as_dict = lambda s: dict(bands=s.bands, branches=s.branches, ms=s.ms, executor=s.executor, order=s.order)  # noqa: E731
    return as_dict(best), [as_dict(tried[i]) for i in range(min(n, 32))]


def _check_srv(rc, srv):
    if rc != OK:
        msg = lib().lux_sgm_server_last_error(srv).decode("utf-8", "replace")
        raise (NotSupported if rc == EUNSUPPORTED else SgmError)(rc, msg)


class Server:
    """The served form (lux_sgm_server_*): requests admitted over `slots` device slots under an HBM budget."""

    def __init__(self, device=0, hbm_budget=0, slots=4, device_loop=False):
        """device_loop: the request loop runs on the device (LUX_SGM_SERVER_DEVICE_LOOP; needs hbm_budget)."""
        h = ctypes.c_void_p()
        flags = SERVER_DEVICE_LOOP if device_loop else 0
        _check_srv(lib().lux_sgm_server_create_ex(device, hbm_budget, slots, flags, ctypes.byref(h)), None)
        self.handle = h
        self._keep = {}

    def submit(self, p, left, right, *, disp=None, sub=None, valid=None):
        """Host arrays (numpy): inputs are copied at once; outputs are written when wait() returns."""
        t = ctypes.c_uint64()
        _check_srv(lib().lux_sgm_server_submit(self.handle, ctypes.byref(p), MEM_HOST, _p(left), _p(right), _p(disp),
                                               _p(sub), _p(valid), ctypes.byref(t)), self.handle)
        self._keep[t.value] = (disp, sub, valid)
        return t.value

    def submit_device(self, p, d_left, d_right, d_disp=0, d_sub=0, d_valid=0):
        """Device pointers (ints) that stay valid until the request completes."""
        t = ctypes.c_uint64()
        vp = lambda x: ctypes.c_void_p(x or None)  # noqa: E731
        _check_srv(lib().lux_sgm_server_submit(self.handle, ctypes.byref(p), MEM_DEVICE, vp(d_left), vp(d_right),
                                               vp(d_disp), vp(d_sub), vp(d_valid), ctypes.byref(t)), self.handle)
        return t.value

    def submit_pyramid(self, pp, left, right, *, disp=None, sub=None):
        """A coarse-to-fine request from host arrays (PyramidParams from pyramid_params)."""
        t = ctypes.c_uint64()
        _check_srv(lib().lux_sgm_server_submit_pyramid(self.handle, ctypes.byref(pp), MEM_HOST, _p(left), _p(right),
                                                       _p(disp), _p(sub), ctypes.byref(t)), self.handle)
        self._keep[t.value] = (disp, sub)
        return t.value

    def submit_pyramid_device(self, pp, d_left, d_right, d_disp=0, d_sub=0):
        t = ctypes.c_uint64()
        vp = lambda x: ctypes.c_void_p(x or None)  # noqa: E731
        _check_srv(lib().lux_sgm_server_submit_pyramid(self.handle, ctypes.byref(pp), MEM_DEVICE, vp(d_left),
                                                       vp(d_right), vp(d_disp), vp(d_sub), ctypes.byref(t)), self.handle)
        return t.value

    def poll(self, ticket):
        st = ctypes.c_int32()
        _check_srv(lib().lux_sgm_server_poll(self.handle, ticket, ctypes.byref(st)), self.handle)
        return st.value

    def wait(self, ticket, timeout_ms=-1):
        """Wait for a request; returns its stats as a dict (raises on a failed request)."""
        st = RequestStats()
        rc = lib().lux_sgm_server_wait(self.handle, ticket, timeout_ms, ctypes.byref(st))
        if rc != ECANCELED:  # on a time-out the request is still queued and will write the outputs
            self._keep.pop(ticket, None)
        _check_srv(rc, self.handle)
        return dict(queued_ms=st.queued_ms, run_ms=st.run_ms, total_ms=st.total_ms, slot=st.slot)

    def cancel(self, ticket):
        _check_srv(lib().lux_sgm_server_cancel(self.handle, ticket), self.handle)

    def info(self):
        i = ServerInfo()
        _check_srv(lib().lux_sgm_server_info_get(self.handle, ctypes.byref(i)), self.handle)
        return {k: getattr(i, k) for k, _ in ServerInfo._fields_ if k != "reserved"}

    def close(self):
        if self.handle:
            lib().lux_sgm_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pyramid_params(W, H, D, radii, **kw):
    """Coarse-to-fine request: full-resolution pair W x H, D planes at the finest level, one radius per finer level."""
    p = PyramidParams()
    p.base = rect_params(W, H, D, **kw)
    p.levels = len(radii) + 1
    for i, r in enumerate(radii):
        p.radius[i] = r
    return p


def pyramid_bytes(p):
    out = ctypes.c_uint64()
    _check(lib().lux_sgm_pyramid_bytes(ctypes.byref(p), ctypes.byref(out)))
    return out.value


def pyramid(ctx, left, right, D, radii, **kw):
    """Coarse-to-fine rectified request from host arrays. Returns dict(disp, sub) of the finest level."""
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    H, W = left.shape
    bits = kw.pop("bits", 8 if left.dtype == np.uint8 else 10)
    p = pyramid_params(W, H, D, radii, bits=bits, **kw)
    disp = np.empty((H, W), np.uint16)
    sub = np.empty((H, W), np.float32)
    _check(lib().lux_sgm_pyramid(ctx.handle, ctypes.byref(p), _p(left), _p(right), _p(disp), _p(sub)), ctx.handle)
    return dict(disp=disp, sub=sub)


def pyramid_device(ctx, p, d_left, d_right, d_disp, d_sub=0, stream=0):
    _check(lib().lux_sgm_pyramid_device(ctx.handle, ctypes.byref(p), ctypes.c_void_p(d_left), ctypes.c_void_p(d_right),
                                        ctypes.c_void_p(d_disp or None), ctypes.c_void_p(d_sub or None),
                                        ctypes.c_void_p(stream or None)), ctx.handle)


def skip_mask(mode, W, H):
    """DenseLayer.SkipMask (DenseStereoSgm.cs:64-87): uint8 [H, W], 0 computed, 0xff skipped."""
    out = np.empty((H, W), np.uint8)
    _check(lib().lux_sgm_skip_mask(mode, W, H, _p(out)))
    return out


def mt19937_draw(seed, n):
    """The n-th output of the 32-bit Mersenne twister seeded with `seed`."""
    return int(lib().lux_sgm_mt19937_draw(seed, n))


class Calib(ctypes.Structure):
    """One camera's CalibData (Triangulator.cs:8-19): K (3x3 row-major), T, R (3x3 row-major)."""
    _fields_ = [("K", ctypes.c_float * 9), ("T", ctypes.c_float * 3), ("R", ctypes.c_float * 9)]


def calib(K, T, R):
    c = Calib()
    c.K[:] = [float(v) for v in np.asarray(K, np.float32).reshape(9)]
    c.T[:] = [float(v) for v in np.asarray(T, np.float32).reshape(3)]
    c.R[:] = [float(v) for v in np.asarray(R, np.float32).reshape(9)]
    return c


def _calibs(cs):
    return (Calib * len(cs))(*[c if isinstance(c, Calib) else calib(*c) for c in cs])


def planes(near_z, far_z, calibs, density, align=1):
    """DenseStereo.Planes (DenseStereo.cs:41-88): float32 plane depths far .. near. calibs: Calib records or (K, T, R)
    tuples, camera 0 first. Raises SgmError with the reference's message where it throws."""
    arr = _calibs(calibs)
    n = ctypes.c_int32(0)
    L = lib()

    def call(out, cap):
        rc = L.lux_sgm_planes(float(near_z), float(far_z), arr, len(arr), float(density), int(align), _p(out), cap,
                              ctypes.byref(n))
        if rc:
            raise SgmError(rc, L.lux_sgm_calib_last_error().decode("utf-8", "replace"))

    call(None, 0)
    out = np.empty(n.value, np.float32)
    call(out, n.value)
    return out


def flow_matrix(a, b):
    """Mat4D.FlowMatrix (ViewTransform.cs:120-133): the column-major float32[16] warp from camera a to camera b."""
    out = np.empty(16, np.float32)
    ca, cb = _calibs([a]), _calibs([b])
    _check(lib().lux_sgm_flow_matrix(ca, cb, _p(out)))
    return out


def homography_sqrt(x):
    """Homography.Sqrt (Homography.cs:18)."""
    return np.float32(lib().lux_sgm_homography_sqrt(float(np.float32(x))))


def rcpss_bits(b):
    return int(lib().lux_sgm_rcpss_bits(int(b) & 0xFFFFFFFF))


def rsqrtss_bits(b):
    return int(lib().lux_sgm_rsqrtss_bits(int(b) & 0xFFFFFFFF))


def box3(ctx, img, bits=None):
    """DenseStereo.BoxFilter3 on the device: uint8/uint16 [H, W] or RGBA8 [H, W, 4]."""
    img = np.ascontiguousarray(img)
    ch = img.shape[2] if img.ndim == 3 else 1
    bits = bits or (8 if img.dtype == np.uint8 else 16)
    out = np.empty_like(img)
    H, W = img.shape[:2]
    _check(lib().lux_sgm_box3(ctx.handle, _p(img), bits, ch, W, H, W, _p(out)), ctx.handle)
    return out


def upsample_layer(ctx, prev_depth, guide, set_w, set_h, sigma=12.0, prev_conf=None):
    """DenseUpsampleLayer.Run (DenseUpsampleLayer.cs:35-48) on the device: float [ph, pw] depth + RGBA8 guide
    [H, W, 4] -> float [H, W] depth, or (depth, confidence) when ``prev_conf`` [ph, pw] is given."""
    prev_depth = np.ascontiguousarray(prev_depth, np.float32)
    guide = np.ascontiguousarray(guide, np.uint8)
    H, W = guide.shape[:2]
    out = np.empty((H, W), np.float32)
    conf = None
    if prev_conf is not None:
        prev_conf = np.ascontiguousarray(prev_conf, np.float32)
        assert prev_conf.shape == prev_depth.shape
        conf = np.empty((H, W), np.float32)
    _check(lib().lux_sgm_upsample_layer(ctx.handle, _p(prev_depth), None if prev_conf is None else _p(prev_conf),
                                        prev_depth.shape[1], prev_depth.shape[0], _p(guide), W, H, set_w, set_h, sigma,
                                        _p(out), None if conf is None else _p(conf)), ctx.handle)
    return out if conf is None else (out, conf)


def upsample_layer_device(ctx, d_prev, pw, ph, d_guide, W, H, set_w, set_h, d_depth, sigma=12.0, d_prev_conf=0, d_conf=0,
                          stream=0):
    """Device-pointer form of upsample_layer: pointers are ints (torch data_ptr()); enqueue only."""
    _check(lib().lux_sgm_upsample_layer_device(ctx.handle, ctypes.c_void_p(d_prev), ctypes.c_void_p(d_prev_conf or None), pw,
                                               ph, ctypes.c_void_p(d_guide), W, H, set_w, set_h, sigma,
                                               ctypes.c_void_p(d_depth), ctypes.c_void_p(d_conf or None),
                                               ctypes.c_void_p(stream or None)), ctx.handle)


class View(ctypes.Structure):
    """lux_sgm_view: one other camera of DenseStereo.CostContext."""
    _fields_ = [("rgba", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("warp", ctypes.c_float * 16), ("wq", ctypes.c_uint16 * 4)]


class Sweep(ctypes.Structure):
    """lux_sgm_sweep: the cost context of a layer (reference, cameras, planes, caps)."""
    _fields_ = [("ref", ctypes.c_void_p), ("ref_width", ctypes.c_int32), ("ref_height", ctypes.c_int32),
                ("n_views", ctypes.c_int32), ("views", ctypes.c_void_p), ("planes", ctypes.c_void_p),
                ("n_planes", ctypes.c_int32), ("caps", ctypes.c_uint8 * 4)]


class CostContext:
    """Mirror of DenseStereo.CostContext (DenseStereo.cs:180-200): ``Ref`` and ``Others`` (box-filtered RGBA8
    [h, w, 4]), ``Warps`` (16 floats each, column-major), ``Wq`` and ``Planes``; holds the arrays the C struct
    points at."""

    def __init__(self, ref, others, warps, wq, planes, caps=(2, 6, 6, 0)):
        self.Ref = np.ascontiguousarray(ref, np.uint8)
        self.Others = [np.ascontiguousarray(o, np.uint8) for o in others]
        self.Warps = [np.ascontiguousarray(w, np.float32).reshape(16) for w in warps]
        self.Wq = [tuple(int(v) for v in q) for q in wq]
        self.Planes = np.ascontiguousarray(planes, np.float32)
        self.Caps = tuple(caps)
        self._views = (View * len(self.Others))()
        for k, (img, w, q) in enumerate(zip(self.Others, self.Warps, self.Wq)):
            v = self._views[k]
            v.rgba = img.ctypes.data
            v.height, v.width = img.shape[:2]
            v.warp = (ctypes.c_float * 16)(*w)
            v.wq = (ctypes.c_uint16 * 4)(*q)
        self.sweep = Sweep(self.Ref.ctypes.data, self.Ref.shape[1], self.Ref.shape[0], len(self.Others),
                           ctypes.addressof(self._views), self.Planes.ctypes.data, len(self.Planes),
                           (ctypes.c_uint8 * 4)(*self.Caps))


def sweep_cost(ctx, cc, W, H, scale, align=8, start=None, count=None, skip=None):
    """The raw volume of a layer (MatchingCost + Normalise per computed pixel), packed: roundup(count, align) bytes
    per pixel in raster order."""
    n = len(cc.Planes)
    if start is not None:
        start = np.ascontiguousarray(start, np.uint16).reshape(-1)
        count = np.ascontiguousarray(count, np.uint16).reshape(-1)
        total = int((((count.astype(np.int64) + align - 1) // align) * align).sum())
    else:
        total = W * H * (-(-n // align) * align)
    skip = None if skip is None else np.ascontiguousarray(skip, np.uint8).reshape(-1)
    raw = np.empty(max(total, 1), np.uint8)
    _check(lib().lux_sgm_sweep_cost(ctx.handle, ctypes.byref(cc.sweep), W, H, scale, align, _p(start), _p(count), _p(skip),
                                    _p(raw)), ctx.handle)
    return raw[:total]


class DenseParams(ctypes.Structure):
    """lux_sgm_dense_params: StereoParams of one layer plus its size."""
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("scale", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("align", ctypes.c_int32), ("p1", ctypes.c_int32), ("p2_scale", ctypes.c_float),
                ("guidance", ctypes.c_float * 4)]


# StereoParams.Guidance (DenseStereo.cs:21): c/6, c/16, c/16, 0 with c = 0x3ef6384f, each one float division
_C_G = np.array([0x3ef6384f], np.uint32).view(np.float32)[0]
L16_GUIDANCE = (float(_C_G / np.float32(6)), float(_C_G / np.float32(16)), float(_C_G / np.float32(16)), 0.0)
# StereoParams.L16Pyramid (DenseStereo.cs:24-31): scales and tiles of the six layers, skip mode 2 on the last two
L16_SCALES = (32, 16, 8, 4, 2, 1)
L16_TILES = (4, 8, 16, 32, 64, 64)


def dense_params(ref_w, ref_h, scale, tile, align=8, p1=1, p2_scale=500.0, guidance=L16_GUIDANCE):
    return DenseParams(-(-ref_w // scale), -(-ref_h // scale), scale, tile, align, p1, p2_scale,
                       (ctypes.c_float * 4)(*guidance))


def range_from_previous(ctx, planes, prev_depth, prev_skip, w, h, margin=1):
    """RangeFromPrevious on the device: (start, count, MinLo, MaxHi) like oracle.range_from_previous."""
    planes = np.ascontiguousarray(planes, np.float32)
    prev_depth = np.ascontiguousarray(prev_depth, np.float32)
    ph, pw = prev_depth.shape
    prev_skip = None if prev_skip is None else np.ascontiguousarray(prev_skip, np.uint8)
    start = np.empty((h, w), np.uint16)
    count = np.empty((h, w), np.uint16)
    lo, hi = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().lux_sgm_range_from_previous(ctx.handle, _p(planes), len(planes), _p(prev_depth), _p(prev_skip), pw, ph, w,
                                             h, margin, _p(start), _p(count), ctypes.byref(lo), ctypes.byref(hi)),
           ctx.handle)
    return start, count, lo.value, hi.value


def dense_layer(ctx, cc, params, start=None, count=None, skip=None, want_agg=False, keep_pass1=False):
    """DenseLayer.Compute on the device from the images: returns dict(depth, wta, agg, agg_pass1) (agg packed)."""
    W, H, align = params.width, params.height, params.align
    n = len(cc.Planes)
    if start is not None:
        start = np.ascontiguousarray(start, np.uint16).reshape(-1)
        count = np.ascontiguousarray(count, np.uint16).reshape(-1)
        total = int((((count.astype(np.int64) + align - 1) // align) * align).sum())
    else:
        total = W * H * (-(-n // align) * align)
    skip = None if skip is None else np.ascontiguousarray(skip, np.uint8).reshape(-1)
    depth = np.empty((H, W), np.float32)
    wta = np.empty((H, W), np.uint16)
    agg = np.empty(max(total, 1), np.uint16) if want_agg else None
    a1 = np.empty(max(total, 1), np.uint16) if keep_pass1 else None
    _check(lib().lux_sgm_dense_layer(ctx.handle, ctypes.byref(cc.sweep), ctypes.byref(params), _p(start), _p(count), _p(skip),
                                     _p(depth), _p(agg), _p(a1), _p(wta)), ctx.handle)
    return dict(depth=depth, wta=wta, agg=None if agg is None else agg[:total],
                agg_pass1=None if a1 is None else a1[:total])


class DensePyramidParams(ctypes.Structure):
    """lux_sgm_dense_pyramid_params: StereoParams.L16Pyramid() and the budget of DenseStereoPyramid."""
    _fields_ = [("levels", ctypes.c_int32), ("scale", ctypes.c_int32 * 8), ("tile", ctypes.c_int32 * 8),
                ("skip_mode", ctypes.c_int32 * 8), ("p1", ctypes.c_int32), ("p2_scale", ctypes.c_float),
                ("guidance", ctypes.c_float * 4), ("margin", ctypes.c_int32), ("caps", ctypes.c_uint8 * 4),
                ("budget", ctypes.c_uint64)]


def dense_pyramid_defaults():
    p = DensePyramidParams()
    lib().lux_sgm_dense_pyramid_defaults(ctypes.byref(p))
    return p


def dense_pyramid(ctx, images, gray, warps, planes8, planes2=None, params=None):
    """DenseStereoPyramid.Run on the device: unfiltered RGBA8 images (reference first), per-camera gray flags, the
    other cameras' warps, the plane lists of align 8 and 2. Returns (list of per-level depth arrays, aligns)."""
    p = params or dense_pyramid_defaults()
    imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
    n = len(imgs)
    ptrs = (ctypes.c_void_p * n)(*[im.ctypes.data for im in imgs])
    ws = (ctypes.c_int32 * n)(*[im.shape[1] for im in imgs])
    hs = (ctypes.c_int32 * n)(*[im.shape[0] for im in imgs])
    g = None if gray is None else np.ascontiguousarray(np.asarray(gray, np.uint8))
    wr = np.ascontiguousarray(np.asarray(warps, np.float32).reshape(n - 1, 16))
    p8 = np.ascontiguousarray(planes8, np.float32)
    p2 = None if planes2 is None else np.ascontiguousarray(planes2, np.float32)
    H0, W0 = imgs[0].shape[:2]
    depths = [np.empty((-(-H0 // p.scale[i]), -(-W0 // p.scale[i])), np.float32) for i in range(p.levels)]
    dptr = (ctypes.c_void_p * p.levels)(*[d.ctypes.data for d in depths])
    aligns = (ctypes.c_int32 * p.levels)()
    _check(lib().lux_sgm_dense_pyramid(ctx.handle, ctypes.byref(p), n, ptrs, ws, hs, _p(g), _p(wr), _p(p8), len(p8), _p(p2),
                                       0 if p2 is None else len(p2), dptr, aligns), ctx.handle)
    return depths, list(aligns)


def dense_pyramid_run(ctx, images, gray, calibs, near=200.0, far=640000.0, density=2.0, params=None):
    """DenseStereoPyramid.Run(images, gray, calibs, near, far) (DenseStereoSgm.cs:331-355) with its set-up: the plane
    lists of align 8 and 2 (DenseStereo.Planes, at the StereoParams plane density, 2 for module type 0 and 4 otherwise)
    and the warps Mat4D.FlowMatrix(calibs[0], calibs[k]) (DenseStereo.BuildContext, :189-199), all from the library's
    host functions. near / far default to DenseStereo.Near / Far. Returns (depths, aligns, planes8, planes2, warps)."""
    p8 = planes(near, far, calibs, density, 8)
    p2 = planes(near, far, calibs, density, 2)
    warps = [flow_matrix(calibs[0], c) for c in calibs[1:]]
    depths, aligns = dense_pyramid(ctx, images, gray, warps, p8, p2, params)
    return depths, aligns, p8, p2, warps


def dense_pyramid_device(ctx, d_images, shapes, gray, warps, planes8, planes2, d_depths, params=None):
    """Device-pointer form: d_images / d_depths are device addresses (ints), shapes [(h, w)] per image."""
    p = params or dense_pyramid_defaults()
    n = len(d_images)
    ptrs = (ctypes.c_void_p * n)(*d_images)
    ws = (ctypes.c_int32 * n)(*[w for _, w in shapes])
    hs = (ctypes.c_int32 * n)(*[h for h, _ in shapes])
    g = None if gray is None else np.ascontiguousarray(np.asarray(gray, np.uint8))
    wr = np.ascontiguousarray(np.asarray(warps, np.float32).reshape(n - 1, 16))
    p8 = np.ascontiguousarray(planes8, np.float32)
    p2 = None if planes2 is None else np.ascontiguousarray(planes2, np.float32)
    dptr = (ctypes.c_void_p * p.levels)(*d_depths)
    aligns = (ctypes.c_int32 * p.levels)()
    _check(lib().lux_sgm_dense_pyramid_device(ctx.handle, ctypes.byref(p), n, ptrs, ws, hs, _p(g), _p(wr), _p(p8), len(p8),
                                              _p(p2), 0 if p2 is None else len(p2), dptr, aligns), ctx.handle)
    return list(aligns)
