"""bench.py's host logic, without a GPU: the one-GPU guard, the kernel-to-stage mapping and the counter summary that
turns two rocprofv3 --pmc runs into the traffic file the bench line reads (its calibration factors, per-request and
per-stage bytes, and the source digest that keeps a stale file out)."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

NS = "void lux::(anonymous namespace)::"
COPY = "__amd_rocclr_copyBuffer"
KERNELS = {  # demangled name: (stage, FETCH_SIZE KiB, WRITE_SIZE KiB) of one dispatch
    NS + "census_kernel<unsigned short, 3>(unsigned short const*, int, int, int, unsigned long*)": (None, 100.0, 50.0),
    NS + "sgm_row_fwd_pv<2, false>(lux::(anonymous namespace)::RowArgs)": ("rows+", 10.0, 400.0),
    NS + "sgm_sweeps_pv<2, 1, false>(lux::(anonymous namespace)::SweepArgs3)": ("vd+", 300.0, 600.0),
    NS + "sgm_sweeps_pv<2, -1, false>(lux::(anonymous namespace)::SweepArgs3)": ("vd-", 310.0, 590.0),
    NS + "sgm_row_bwd_wta_pv<2, false, 0>(lux::(anonymous namespace)::RowArgs)": ("rows-", 700.0, 5.0),
    NS + "sgm_replay_chain<2, false, false, false>(lux::(anonymous namespace)::ReplayArgs, int const*)": (None, 1.0, 1.0),
}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("name,stage", [(k, v[0]) for k, v in KERNELS.items()] +
                         [(NS + "sgm_sweep<2, -1, 3, false, false, false, false>(x)", "vd-"),
                          (NS + "sgm_row_bwd_wta<2, false, false, false>(x)", "rows-"), (COPY, None)])
def test_stage_of_kernel_names(name, stage):
    assert bench._stage_of(name) == stage


def _write_pmc(d, counter, requests, copy_kib):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "pmc_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        n = 0
        for _ in range(requests):
            for k, (_, f, wr) in KERNELS.items():
                for _ in range(2 if "census_kernel" in k else 1):  # a request's two census launches
                    n += 1
                    w.writerow([n, k, counter, f if counter == "FETCH_SIZE" else wr])
        n += 1
        w.writerow([n, COPY, counter, copy_kib])


def test_summarize_pmc_factors_and_stages(tmp_path, monkeypatch):
    out = tmp_path / "traffic.json"
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(out))
    # the calibration copy moves 1 GiB: FETCH_SIZE reads half of it (factor 2), WRITE_SIZE all of it (factor 1)
    _write_pmc(str(tmp_path / "fetch"), "FETCH_SIZE", 3, (1 << 30) / 1024 / 2)
    _write_pmc(str(tmp_path / "write"), "WRITE_SIZE", 3, (1 << 30) / 1024)
    bench.summarize_pmc(str(tmp_path / "fetch"), str(tmp_path / "write"))
    t = json.loads(out.read_text())
    assert t["read_factor"] == pytest.approx(2.0) and t["write_factor"] == pytest.approx(1.0)
    assert t["requests"] == 3
    per = {k: (2 if "census" in k else 1) * (2.0 * f + wr) * 1024 for k, (_, f, wr) in KERNELS.items()}
    assert t["request_bytes"] == pytest.approx(sum(per.values()), abs=8)
    for k, (st, _, _) in KERNELS.items():
        if st:
            assert t["stages"][st] == pytest.approx(per[k], abs=1)
    assert COPY not in str(t["kernels"])
    # read back only while the digest matches the sources
    data, note = bench.load_traffic()
    assert data is not None and data["request_bytes"] == t["request_bytes"], note
    t["source_sha16"] = "0" * 16
    out.write_text(json.dumps(t))
    data, note = bench.load_traffic()
    assert data is None and "not used" in note

