"""Host logic of bench.py: the roofline's traffic comes from the committed
rocprofv3 summary only for the workload that summary measured, and a run on
more than one GPU is refused with a message."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_of_the_profiled_workload_only():
    kv = None
    for line in open(bench.PMC_FILE):
        if line.startswith("traffic k_pp<float"):
            t = line.split(">", 1)[1].split()
            kv = dict(zip(t[0::2], map(float, t[1::2])))
    assert kv is not None
    N, M, rc = int(kv["N"]), int(kv["M"]), kv["rc"]
    t = bench.pmc_traffic("pp", "mixed", N, M, rc, kv["alpha"])
    expect = (kv["fetch_kib"] * kv["fetch_factor"] + kv["write_kib"] * kv["write_factor"]) * 1024.0
    assert t == expect and t > 56e6  # at least the algorithmic 56 B a particle
    assert bench.pmc_traffic("pp", "mixed", N, M, rc + 0.25, kv["alpha"]) is None
    assert bench.pmc_traffic("pp", "fp64", N, M, rc, kv["alpha"]) is None
    assert bench.pmc_traffic("gather", "mixed", N, M, rc, kv["alpha"]) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
