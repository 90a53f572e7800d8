"""Host-side parts of bench.py, on the CPU: the refusal of runs over several
GPUs, the time-bounded repetition behind the CPU baseline, the lookup of
measured HBM traffic, and the CPU baseline itself on small workloads."""
import json
import os
import subprocess
import sys

import pytest

import bench
from dedispersion_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("gpus", [2, 8])
def test_several_gpus_refused(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(gpus)],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == gpus
    assert "one GPU" in line["error"]


def test_repeat_for_runs_at_least_once_and_stops_at_the_cap():
    calls = []
    passes, dt = bench.repeat_for(lambda: calls.append(1), 0.0)
    assert passes == 1 and len(calls) == 1 and dt >= 0.0
    calls.clear()
    passes, _ = bench.repeat_for(lambda: calls.append(1), 3600.0, max_passes=7)
    assert passes == 7 and len(calls) == 7


def test_measured_traffic_matches_only_the_profiled_conf():
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
        rec = json.load(f)
    name = "batched-beams-40x1536ch-two-step"
    for kernel in ("step_one_batch", "step_two_batch"):
        want = rec[name][kernel]
        assert bench.measured_traffic(name, kernel, want["conf"]) == want["traffic_bytes"]
        assert bench.measured_traffic(name, kernel, "0 1 2 32 2 1 8 4 0") is None
    assert bench.measured_traffic(name, "no_such_kernel", "0 1 4 64 4 1 16 8 1") is None
    assert bench.measured_traffic("no-such-workload", "step_two_batch", "0 1 4 64 4 1 16 8 1") is None


@pytest.mark.parametrize("mode,ws", [
    ("single", [workloads.tiny(channels=32, samples=256, dms=16, step=1.0, zapped=(3,))]),
    ("two-step", [workloads.tiny(channels=32, samples=256, dms=4, step=0.5, subbands=4, coarse=4, coarse_step=2.0),
                  workloads.tiny(channels=32, samples=512, dms=4, step=0.5, subbands=4, coarse=8, coarse_step=2.0,
                                 seed=6)]),
])
def test_cpu_baseline_record(mode, ws):
    rec = bench.cpu_baseline(ws, mode, 0.2)
    assert rec["kind"] == "port" and rec["unit"] == "DM-trial*samples/s"
    host = rec["host"]
    assert rec["cores"] == host["threads"] and 1 <= host["threads"] <= host["affinity_cpus"]
    assert host["model"] and host["logical_cpus"] >= host["affinity_cpus"]
    assert rec["value"] > 0.0
    if mode == "two-step":
        assert "of each of the 2 beams of the mix" in rec["sample"]
    else:
        assert ws[0].name in rec["sample"] and "pass(es)" in rec["sample"]
    one = rec["single_thread"]
    assert one["cores"] == 1 and one["value"] > 0.0


def test_host_cpu_threads_follow_the_process_share(monkeypatch):
    """The baseline uses every CPU the process may run on, or the per-process
    share the environment sets when that is smaller."""
    affinity = len(os.sched_getaffinity(0))
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    assert bench.host_cpu()["threads"] == affinity
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_cpu()["threads"] == 1
    monkeypatch.setenv("OMP_NUM_THREADS", str(affinity + 100))
    assert bench.host_cpu()["threads"] == affinity
