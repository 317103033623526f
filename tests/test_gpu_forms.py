"""The batched (one graph launch) and served (request loop) forms: every result equals the plain drop-in's on
the same pair alone, bit for bit, whatever the mix, schedule or admission order; the plain drop-in itself is
checked against the oracle in test_gpu_parity.py and once more here."""
import numpy as np
import pytest

from lux_amd import scene, sgm
from oracle import oracle as o

pytestmark = pytest.mark.gpu

# (W, H, D, bits, cost, radius, seed): mixed sizes, disparity ranges, pixel types and costs
MIX = [
    (96, 64, 32, 8, "census", 2, 1),
    (160, 120, 64, 10, "census", 3, 2),
    (130, 70, 48, 8, "ad", 0, 3),
    (200, 100, 96, 10, "census", 3, 4),
    (64, 48, 16, 8, "census", 1, 5),
    (250, 90, 128, 10, "census", 3, 6),
    (77, 33, 24, 12, "ad", 0, 7),
]


def _scene(W, H, D, bits, seed):
    left, right, _ = scene.make_pair(W, H, max(D - 1, 1), planes=4, sigma=2.0, seed=seed, bits=bits)
    return left, right


def _params(W, H, D, bits, cost, radius, tile=16):
    return sgm.rect_params(W, H, D, bits=bits, cost=cost, radius=max(radius, 1), ad_cap=20, p1=8, p2=96, tile=tile)


@pytest.fixture(scope="module")
def cases():
    out = []
    with sgm.Context(0) as c:
        for (W, H, D, bits, cost, radius, seed) in MIX:
            left, right = _scene(W, H, D, bits, seed)
            p = _params(W, H, D, bits, cost, radius)
            r = sgm.rectified(c, left, right, D, bits=bits, cost=cost, radius=max(radius, 1), ad_cap=20, p1=8, p2=96,
                              tile=16)
            out.append(dict(W=W, H=H, D=D, bits=bits, cost=cost, left=left, right=right, p=p, alone=r))
    return out


def test_alone_matches_oracle(cases):
    c = cases[0]
    left, right = c["left"], c["right"]
    ref = o.rectified(left, right, c["D"], bits=8, cost="census", radius=2, p1=8, p2=96, T=16, align=8, tol=1)
    np.testing.assert_array_equal(c["alone"]["disp"], ref["disp"])
    np.testing.assert_array_equal(c["alone"]["valid"], ref["valid"])
    np.testing.assert_allclose(c["alone"]["sub"], ref["sub"], rtol=0, atol=1e-4)


def _device_pairs(torch, cases):
    dev = torch.device("cuda:0")
    pairs, outs, keep = [], [], []
    for c in cases:
        dt = torch.uint8 if c["bits"] == 8 else torch.int16
        L = torch.from_numpy(c["left"].view(np.uint8 if c["bits"] == 8 else np.int16)).to(dev)
        R = torch.from_numpy(c["right"].view(np.uint8 if c["bits"] == 8 else np.int16)).to(dev)
        assert L.dtype == dt
        d = torch.zeros((c["H"], c["W"]), dtype=torch.int16, device=dev)
        s = torch.zeros((c["H"], c["W"]), dtype=torch.float32, device=dev)
        v = torch.zeros((c["H"], c["W"]), dtype=torch.uint8, device=dev)
        pr = sgm.Pair()
        pr.params = c["p"]
        pr.left, pr.right, pr.disp, pr.sub, pr.valid = L.data_ptr(), R.data_ptr(), d.data_ptr(), s.data_ptr(), v.data_ptr()
        pairs.append(pr)
        outs.append((d, s, v))
        keep += [L, R]
    return pairs, outs, keep


def _check_outs(outs, cases, what):
    for (d, s, v), c in zip(outs, cases):
        tag = f"{what} {c['W']}x{c['H']} D{c['D']}"
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint16), c["alone"]["disp"], err_msg=tag)
        np.testing.assert_array_equal(s.cpu().numpy(), c["alone"]["sub"], err_msg=tag)
        np.testing.assert_array_equal(v.cpu().numpy(), c["alone"]["valid"], err_msg=tag)


def test_batch_graph_replays(cases):
    torch = pytest.importorskip("torch")
    with sgm.Context(0) as ctx:
        pairs, outs, keep = _device_pairs(torch, cases)
        stream = torch.cuda.Stream()
        for rep in range(3):  # replays of one cached graph, with eager work and syncs between
            print(f"batch replay {rep}", flush=True)
            for d, s, v in outs:
                d.fill_(-1)
                s.fill_(-7.0)
                v.fill_(9)
            torch.cuda.synchronize()
            sgm.batch(ctx, pairs, stream=stream.cuda_stream)
            stream.synchronize()
            _check_outs(outs, cases, f"batch replay {rep}")
        del keep


def test_batch_tuned_schedules(cases):
    torch = pytest.importorskip("torch")
    with sgm.Context(0) as ctx:
        pairs, outs, keep = _device_pairs(torch, cases)
        stream = torch.cuda.Stream()
        best, tried = sgm.batch_tune(ctx, pairs, stream=stream.cuda_stream, reps=1)
        print("tuned", best, tried, flush=True)
        assert len(tried) >= 4 and best["ms"] > 0
        assert {t["bands"] for t in tried} == {0, 1}  # every pair fits the band kernel (cap <= 256)
        _check_outs(outs, cases, "after tuning")  # the last candidate's run
        for d, s, v in outs:
            d.fill_(-1)
        torch.cuda.synchronize()
        sgm.batch(ctx, pairs, stream=stream.cuda_stream)  # the tuned schedule
        stream.synchronize()
        _check_outs(outs, cases, "tuned batch")
        del keep


@pytest.mark.parametrize("executor,order", [(1, 0), (1, 1), (7, 0), (7, 1)])
def test_executor_batch_matches_alone(cases, executor, order):
    """The persistent executor (one launch, or one per pass group; both item orders): every census pair of the mix,
    replayed three times, equals the pair alone bit for bit."""
    torch = pytest.importorskip("torch")
    sel = [c for c in cases if c["cost"] == "census"]
    with sgm.Context(0) as ctx:
        pairs, outs, keep = _device_pairs(torch, sel)
        sgm.batch_set_schedule(ctx, pairs, executor=executor, order=order)
        stream = torch.cuda.Stream()
        for rep in range(3):
            print(f"executor {executor} order {order} run {rep}", flush=True)
            for d, s, v in outs:
                d.fill_(-1)
                s.fill_(-7.0)
                v.fill_(9)
            torch.cuda.synchronize()
            sgm.batch(ctx, pairs, stream=stream.cuda_stream)
            stream.synchronize()
            _check_outs(outs, sel, f"executor {executor}/{order} run {rep}")
        del keep


def test_executor_refuses_other_pairs(cases):
    """AD pairs (and any pair whose path costs do not fit a byte) stay on the graph branches."""
    torch = pytest.importorskip("torch")
    with sgm.Context(0) as ctx:
        pairs, outs, keep = _device_pairs(torch, cases)
        with pytest.raises(sgm.NotSupported):
            sgm.batch_set_schedule(ctx, pairs, executor=1)
        del keep


def _submit_all(srv, cases, order):
    tickets = {}
    for i in order:
        c = cases[i]
        outs = (np.full((c["H"], c["W"]), 0xffff, np.uint16), np.full((c["H"], c["W"]), -1, np.float32),
                np.full((c["H"], c["W"]), 7, np.uint8))
        t = srv.submit(c["p"], c["left"], c["right"], disp=outs[0], sub=outs[1], valid=outs[2])
        tickets[t] = (i, outs)
    return tickets


def _collect(srv, tickets, cases):
    stats = []
    for t, (i, (d, s, v)) in tickets.items():
        st = srv.wait(t, timeout_ms=60000)
        stats.append(st)
        c = cases[i]
        tag = f"served {c['W']}x{c['H']} D{c['D']}"
        np.testing.assert_array_equal(d, c["alone"]["disp"], err_msg=tag)
        np.testing.assert_array_equal(s, c["alone"]["sub"], err_msg=tag)
        np.testing.assert_array_equal(v, c["alone"]["valid"], err_msg=tag)
    return stats


def test_server_mix_equals_alone(cases):
    order = list(range(len(cases))) * 3
    np.random.default_rng(0).shuffle(order)
    with sgm.Server(0, 0, slots=4) as srv:
        print(f"served mix of {len(order)} requests over 4 slots", flush=True)
        tickets = _submit_all(srv, cases, order)
        stats = _collect(srv, tickets, cases)
        info = srv.info()
        assert info["completed"] == len(order) and info["failed"] == 0
        assert len({st["slot"] for st in stats}) > 1  # the work was spread over slots
        assert all(st["total_ms"] >= st["run_ms"] >= 0 for st in stats)


def test_server_device_memory(cases):
    torch = pytest.importorskip("torch")
    pairs, outs, keep = _device_pairs(torch, cases)
    torch.cuda.synchronize()
    with sgm.Server(0, 0, slots=3) as srv:
        ts = [srv.submit_device(pr.params, pr.left, pr.right, pr.disp, pr.sub, pr.valid) for pr in pairs]
        for t in ts:
            srv.wait(t, timeout_ms=60000)
    _check_outs(outs, cases, "served device")
    del keep


def test_server_budget_admission(cases):
    big = cases[5]  # the largest workspace of the mix
    need = sgm.rect_bytes(big["W"], big["H"], big["D"], bits=big["bits"], tile=16)
    budget = int(need * 1.5)  # room for one large request and some small ones at a time
    with sgm.Server(0, budget, slots=4) as srv:
        order = [5, 3, 5, 1, 5, 0, 3]
        tickets = _submit_all(srv, cases, order)
        _collect(srv, tickets, cases)
        info = srv.info()
        assert info["bytes_peak"] <= budget
        assert info["completed"] == len(order)
        huge = sgm.rect_params(4000, 3000, 256, bits=10)
        with pytest.raises(sgm.SgmError) as e:
            srv.submit(huge, np.zeros((3000, 4000), np.uint16), np.zeros((3000, 4000), np.uint16))
        assert e.value.code == sgm.ENOMEM


def test_server_cancel_pending(cases):
    with sgm.Server(0, 0, slots=1) as srv:
        c = cases[5]
        ts = [srv.submit(c["p"], c["left"], c["right"]) for _ in range(6)]
        last = ts[-1]
        try:
            srv.cancel(last)
            cancelled = True
        except sgm.SgmError:
            cancelled = False  # it was admitted already
        for t in ts[:-1]:
            srv.wait(t, timeout_ms=60000)
        if cancelled:
            with pytest.raises(sgm.SgmError) as e:
                srv.wait(last, timeout_ms=60000)
            assert e.value.code == sgm.ECANCELED
        else:
            srv.wait(last, timeout_ms=60000)


def _census_cases(cases):
    return [i for i, c in enumerate(cases) if c["cost"] == "census"]


def test_device_loop_mix_equals_alone(cases):
    """The device-resident loop: the census requests of the mix, three times each in a shuffled order, every result
    equal to the request alone, with several slots in use at once."""
    order = _census_cases(cases) * 3
    np.random.default_rng(1).shuffle(order)
    with sgm.Server(0, 4 << 30, slots=4, device_loop=True) as srv:
        print(f"device loop: mix of {len(order)} requests over 4 slots", flush=True)
        tickets = _submit_all(srv, cases, order)
        stats = _collect(srv, tickets, cases)
        info = srv.info()
        assert info["completed"] == len(order) and info["failed"] == 0
        assert 0 < info["bytes_peak"] <= 4 << 30
        assert len({st["slot"] for st in stats}) > 1
        assert all(st["total_ms"] >= st["run_ms"] >= 0 for st in stats)


def test_device_loop_device_memory_and_relaunch(cases):
    """Device buffers in and out, and a second round after the loop has left for lack of work."""
    import time

    torch = pytest.importorskip("torch")
    sel = [cases[i] for i in _census_cases(cases)]
    pairs, outs, keep = _device_pairs(torch, sel)
    torch.cuda.synchronize()
    with sgm.Server(0, 4 << 30, slots=2, device_loop=True) as srv:
        for rnd in range(2):
            print(f"device loop, device memory, round {rnd}", flush=True)
            for d, s, v in outs:
                d.fill_(-1)
                s.fill_(-7.0)
                v.fill_(9)
            torch.cuda.synchronize()
            ts = [srv.submit_device(pr.params, pr.left, pr.right, pr.disp, pr.sub, pr.valid) for pr in pairs]
            for t in ts:
                srv.wait(t, timeout_ms=60000)
            _check_outs(outs, sel, f"device loop round {rnd}")
            time.sleep(0.1)  # longer than the loop's idle time: the next round launches it again
    del keep


def test_device_loop_refusals(cases):
    """What the device loop does not take fails at submit; a request larger than the pool too."""
    c = cases[5]
    with sgm.Server(0, 1 << 30, slots=2, device_loop=True) as srv:
        ad = cases[2]
        assert ad["cost"] == "ad"
        with pytest.raises(sgm.NotSupported):
            srv.submit(ad["p"], ad["left"], ad["right"])
        with pytest.raises(sgm.SgmError) as e:
            huge = sgm.rect_params(2000, 1500, 128, bits=10, cost="census", radius=3, p1=8, p2=96, tile=16)
            srv.submit(huge, np.zeros((1500, 2000), np.uint16), np.zeros((1500, 2000), np.uint16))
        assert e.value.code == sgm.ENOMEM
        t = srv.submit(c["p"], c["left"], c["right"])  # still serving after the refusals
        srv.wait(t, timeout_ms=60000)
    with pytest.raises(sgm.SgmError):
        sgm.Server(0, 0, slots=2, device_loop=True)  # the pool needs a budget


def test_server_pyramids_mixed_with_pairs(cases):
    """Coarse-to-fine requests in the same queue as plain pairs: each equals the same request run alone."""
    pyr = []
    for i, (W, H, D, radii) in enumerate([(160, 120, 64, (8, 4)), (130, 66, 64, (8,)), (96, 80, 32, (4, 4))]):
        left, right, _ = scene.make_pair(W, H, D - 1, planes=4, sigma=2.0, seed=50 + i, bits=10, fg_frac=0.1)
        with sgm.Context(0) as c:
            alone = sgm.pyramid(c, left, right, D, radii, bits=10, p1=8, p2=96, tile=16)
        pp = sgm.pyramid_params(W, H, D, radii, bits=10, p1=8, p2=96, tile=16)
        pyr.append((pp, left, right, alone))
    with sgm.Server(0, 0, slots=3) as srv:
        tickets, outs = [], []
        for rep in range(2):
            for k, (pp, left, right, alone) in enumerate(pyr):
                d = np.full(left.shape, 0xffff, np.uint16)
                s = np.full(left.shape, -1, np.float32)
                tickets.append(srv.submit_pyramid(pp, left, right, disp=d, sub=s))
                outs.append((k, d, s))
                c = cases[(rep + k) % len(cases)]
                tickets.append(srv.submit(c["p"], c["left"], c["right"]))
                outs.append(None)
        for t, o_ in zip(tickets, outs):
            srv.wait(t, timeout_ms=60000)
            if o_ is not None:
                k, d, s = o_
                np.testing.assert_array_equal(d, pyr[k][3]["disp"])
                np.testing.assert_array_equal(s, pyr[k][3]["sub"])
        assert srv.info()["failed"] == 0


def test_pyramid_without_path_volumes():
    """A budget of exactly the pyramid's own buffers leaves no room for the three path volumes: the ranged levels then
    run their paths one launch each (the fallback), with the results of the one-launch passes."""
    W, H, D, radii = 160, 120, 64, (8, 4)
    left, right, _ = scene.make_pair(W, H, D - 1, planes=4, sigma=2.0, seed=77, bits=10, fg_frac=0.1)
    pp = sgm.pyramid_params(W, H, D, radii, bits=10, p1=8, p2=96, tile=16)
    need = sgm.pyramid_bytes(pp)
    print(f"pyramid {W}x{H} D{D} radii {radii} budget {need}", flush=True)
    with sgm.Context(0) as c:
        full = sgm.pyramid(c, left, right, D, radii, bits=10, p1=8, p2=96, tile=16)
    with sgm.Context(0, hbm_budget=need) as c:
        tight = sgm.pyramid(c, left, right, D, radii, bits=10, p1=8, p2=96, tile=16)
    np.testing.assert_array_equal(tight["disp"], full["disp"])
    np.testing.assert_array_equal(tight["sub"], full["sub"])
