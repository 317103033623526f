"""The workloads at their stated sizes, through the C ABI, checked against the oracle bit for bit:

* full module: one 4160x3120 uint16 10-bit pair, census 7x7, D=256, P1=8, P2=96, T=64: aggregate, disparity,
  subpixel and validity equal the oracle's (DenseStereoSgm.cs:181-232 plus the census front end, WTA, parabola
  and L/R check);
* batched capture: 6 x 2080x1560 D=128 + 4 x 3120x2340 D=192 in one graph launch under the autotuned schedule:
  every pair equals itself run alone, and one pair of each size equals the oracle;
* served mix: one request of each size 640x480 D=64, 1040x780 D=96, 2080x1560 D=128, 4160x3120 D=256 through the
  request loop under a 96 GB budget, the host dispatcher's and the device-resident one: each equals itself run
  alone, and the 4160x3120 and 640x480 requests equal the oracle.

The oracle runs on the host (its aggregation serial, as the reference's, DenseStereoSgm.cs:171); the full-module
case takes it about half a minute.
"""
import os

import numpy as np
import pytest

from lux_amd import scene, sgm
from oracle import oracle as o

pytestmark = pytest.mark.gpu
SUB_TOL = 1e-4  # px
THREADS = max(1, min(16, os.cpu_count() or 1))
KW = dict(bits=10, cost="census", radius=3, p1=8, p2=96, tile=64)


def _say(msg):
    print(msg, flush=True)


def _oracle(left, right, D, want_agg=False):
    return o.rectified(left, right, D, cost="census", radius=3, bits=10, p1=8, p2=96, T=64, threads=THREADS,
                       want_agg=want_agg)


def _same(got, want, tag, agg=False):
    np.testing.assert_array_equal(got["disp"], want["disp"], err_msg=tag)
    np.testing.assert_array_equal(got["valid"], want["valid"], err_msg=tag)
    assert np.abs(got["sub"] - want["sub"]).max() <= SUB_TOL, tag
    np.testing.assert_array_equal(got["sub"], want["sub"], err_msg=tag)
    if agg:
        np.testing.assert_array_equal(got["agg"], want["agg"], err_msg=tag)


@pytest.fixture(scope="module")
def full_pair():
    _say("scene 4160x3120 D=256 (6 planes, sigma 4, seed 1)")
    left, right, _ = scene.make_pair(4160, 3120, 255, planes=6, sigma=4.0, seed=1, bits=10)
    return left, right


@pytest.fixture(scope="module")
def full_oracle(full_pair):
    left, right = full_pair
    _say(f"oracle 4160x3120 D=256 on {THREADS} threads")
    return _oracle(left, right, 256, want_agg=True)


def test_full_module_matches_oracle(full_pair, full_oracle):
    left, right = full_pair
    _say("device 4160x3120 D=256 with the aggregate")
    with sgm.Context(0) as ctx:
        got = sgm.rectified(ctx, left, right, 256, want_agg=True, **KW)
    _same(got, full_oracle, "4160x3120 D=256", agg=True)
    assert got["valid"].mean() > 0.8


SERVED = [(640, 480, 64, 11), (1040, 780, 96, 12), (2080, 1560, 128, 13), (4160, 3120, 256, None)]


def test_served_sizes_match_alone_and_oracle(full_pair, full_oracle):
    reqs = []
    for W, H, D, seed in SERVED:
        if seed is None:
            left, right = full_pair
        else:
            left, right, _ = scene.make_pair(W, H, D - 1, planes=6, sigma=4.0, seed=seed, bits=10)
        reqs.append((W, H, D, left, right))
    with sgm.Context(0) as ctx:
        alone = []
        for W, H, D, left, right in reqs:
            _say(f"alone {W}x{H} D={D}")
            alone.append(sgm.rectified(ctx, left, right, D, **KW))
    for loop in (False, True):  # the host dispatcher, then the device-resident loop
        _serve_four(reqs, alone, loop)
    _same(alone[3], full_oracle, "4160x3120 D=256 served/alone vs oracle")
    W, H, D, left, right = reqs[0]
    _say(f"oracle {W}x{H} D={D}")
    _same(alone[0], _oracle(left, right, D), f"{W}x{H} D={D} vs oracle")


def _serve_four(reqs, alone, device_loop):
    _say(f"served{' (device loop)' if device_loop else ''}: the four sizes, twice, over 4 slots under a 96 GB budget")
    with sgm.Server(0, 96 << 30, slots=4, device_loop=device_loop) as srv:
        order = [3, 0, 2, 1, 0, 3, 1, 2]
        tickets = []
        for i in order:
            W, H, D, left, right = reqs[i]
            outs = dict(disp=np.full((H, W), 0xffff, np.uint16), sub=np.full((H, W), -1, np.float32),
                        valid=np.full((H, W), 7, np.uint8))
            p = sgm.rect_params(W, H, D, **KW)
            tickets.append((srv.submit(p, left, right, **outs), i, outs))
        for t, i, outs in tickets:
            srv.wait(t, timeout_ms=120000)
            W, H, D = reqs[i][:3]
            _same(outs, alone[i], f"served {W}x{H} D={D} vs alone")
        info = srv.info()
        assert info["completed"] == len(order) and info["failed"] == 0
        assert info["bytes_peak"] <= 96 << 30


BATCH = [(2080, 1560, 128)] * 6 + [(3120, 2340, 192)] * 4


def test_batched_capture_matches_alone_and_oracle():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    scenes = []
    for k, (W, H, D) in enumerate(BATCH):
        _say(f"scene {k}: {W}x{H} D={D} seed {100 + k}")
        left, right, _ = scene.make_pair(W, H, D - 1, planes=6, sigma=4.0, seed=100 + k, bits=10)
        scenes.append((W, H, D, left, right))
    with sgm.Context(0) as ctx:
        alone = []
        for W, H, D, left, right in scenes:
            alone.append(sgm.rectified(ctx, left, right, D, **KW))
        pairs, outs, keep = [], [], []
        for W, H, D, left, right in scenes:
            L = torch.from_numpy(left.view(np.int16)).to(dev)
            R = torch.from_numpy(right.view(np.int16)).to(dev)
            d = torch.full((H, W), -1, dtype=torch.int16, device=dev)
            s = torch.full((H, W), -7.0, dtype=torch.float32, device=dev)
            v = torch.full((H, W), 9, dtype=torch.uint8, device=dev)
            pr = sgm.Pair()
            pr.params = sgm.rect_params(W, H, D, **KW)
            pr.left, pr.right = L.data_ptr(), R.data_ptr()
            pr.disp, pr.sub, pr.valid = d.data_ptr(), s.data_ptr(), v.data_ptr()
            pairs.append(pr)
            outs.append((d, s, v))
            keep += [L, R]
        torch.cuda.synchronize()
        stream = torch.cuda.Stream()
        _say("autotuning the 10-pair batch")
        best, tried = sgm.batch_tune(ctx, pairs, stream=stream.cuda_stream, reps=1)
        _say(f"tuned: {best} of {len(tried)} candidates")
        for rep in range(2):
            for d, s, v in outs:
                d.fill_(-1)
                s.fill_(-7.0)
                v.fill_(9)
            torch.cuda.synchronize()
            _say(f"batch launch {rep}")
            sgm.batch(ctx, pairs, stream=stream.cuda_stream)
            stream.synchronize()
            for k, ((d, s, v), a) in enumerate(zip(outs, alone)):
                got = dict(disp=d.cpu().numpy().view(np.uint16), sub=s.cpu().numpy(), valid=v.cpu().numpy())
                _same(got, a, f"batch {rep} pair {k} vs alone")
        for executor, order in ((1, 1), (7, 1)):  # the persistent executor at this size, whatever the tuner chose
            sgm.batch_set_schedule(ctx, pairs, executor=executor, order=order)
            for d, s, v in outs:
                d.fill_(-1)
                s.fill_(-7.0)
                v.fill_(9)
            torch.cuda.synchronize()
            _say(f"executor {executor} order {order}")
            sgm.batch(ctx, pairs, stream=stream.cuda_stream)
            stream.synchronize()
            for k, ((d, s, v), a) in enumerate(zip(outs, alone)):
                got = dict(disp=d.cpu().numpy().view(np.uint16), sub=s.cpu().numpy(), valid=v.cpu().numpy())
                _same(got, a, f"executor {executor}/{order} pair {k} vs alone")
        del keep
    for k in (0, 6):  # one pair of each size against the oracle
        W, H, D, left, right = scenes[k]
        _say(f"oracle pair {k}: {W}x{H} D={D}")
        _same(alone[k], _oracle(left, right, D), f"pair {k} {W}x{H} D={D} vs oracle")
