"""lux_sgm_rectified_device_noalloc: the rectified request with its buffers in the library's static scratch equals
lux_sgm_rectified_device on the same inputs bit for bit (aggregate, disparity, subpixel, validity), on two calls in a
row, the second smaller than the first (the scratch keeps the first call's contents), and refuses a request larger than
the scratch with LUX_SGM_ENOMEM."""
import numpy as np
import pytest

from lux_amd import scene, sgm

pytestmark = pytest.mark.gpu


def _run(torch, ctx, left, right, p, fn):
    dev = torch.device("cuda:0")
    H, W = left.shape
    view = np.uint8 if left.dtype == np.uint8 else np.int16
    L = torch.from_numpy(left.view(view)).to(dev)
    R = torch.from_numpy(right.view(view)).to(dev)
    cap = ((p.disparities + p.align - 1) // p.align) * p.align
    d = torch.full((H, W), -1, dtype=torch.int16, device=dev)
    s = torch.full((H, W), -7.0, dtype=torch.float32, device=dev)
    v = torch.full((H, W), 9, dtype=torch.uint8, device=dev)
    a = torch.full((H, W, cap), -1, dtype=torch.int16, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    fn(ctx, p, L.data_ptr(), R.data_ptr(), d.data_ptr(), s.data_ptr(), v.data_ptr(), a.data_ptr(), stream)
    torch.cuda.synchronize()
    return dict(disp=d.cpu().numpy().view(np.uint16), sub=s.cpu().numpy(), valid=v.cpu().numpy(),
                agg=a.cpu().numpy().view(np.uint16))


def _same(got, want, tag):
    for k in ("agg", "disp", "valid", "sub"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{tag}: {k}")


def test_noalloc_equals_pooled_entry_point():
    torch = pytest.importorskip("torch")
    cases = [
        # 640x480 uint8, census 5x5, D=64, P1=10, P2=120 (the CPU reference path's workload), then a smaller uint16 pair
        (640, 480, 64, 8, 2, 10, 120, 0),
        (200, 120, 48, 10, 3, 8, 96, 1),
    ]
    with sgm.Context(0) as ctx:
        for W, H, D, bits, radius, p1, p2, seed in cases:
            left, right, _ = scene.make_pair(W, H, D - 1, planes=4, sigma=2.0, seed=seed, bits=bits)
            p = sgm.rect_params(W, H, D, bits=bits, cost="census", radius=radius, p1=p1, p2=p2, tile=64)
            print(f"no-alloc {W}x{H} D={D} bits={bits} census r={radius}", flush=True)
            want = _run(torch, ctx, left, right, p, sgm.rectified_device)
            got = _run(torch, ctx, left, right, p, sgm.rectified_device_noalloc)
            _same(got, want, f"{W}x{H} D={D}")
            assert (want["valid"] > 0).mean() > 0.5
        W, H, D = 1040, 780, 96
        p = sgm.rect_params(W, H, D, bits=10, cost="census", radius=3, p1=8, p2=96, tile=64)
        with pytest.raises(sgm.SgmError) as e:  # refused on the host: nothing is enqueued
            sgm.rectified_device_noalloc(ctx, p, 1, 1, 0, 0, 0)
        assert e.value.code == sgm.ENOMEM
