"""bench.py's command-line contract on the CPU: one GPU per run (any other
--gpus is a JSON error line and a non-zero exit, not a silent fallback), and
the helpers that turn the committed counter summaries into roofline traffic."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused_with_a_json_error():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=str(ROOT))
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "error" in line and line["n_gpus"] == 2


def test_traffic_entries_resolve_for_the_headline_and_graph_kernels():
    sys.path.insert(0, str(ROOT))
    try:
        import bench
    finally:
        sys.path.pop(0)
    # the headline int8 kernel (11 moduli, configuration 944) and the gapped
    # water-box FP64 batched kernel have committed FETCH / WRITE summaries
    t = bench.traffic_bytes("int8crt", 8192, 11, 944)
    assert t is not None and 1e9 < t < 2e10
    assert bench.traffic_bytes("graph", 30000, None, "gapped") is not None
    # a configuration without a summary gives no traffic figure
    assert bench.traffic_bytes("int8crt", 8192, 11, 803) is None
    assert bench.crt_kernel_name(944) == "crt_product_persistent_kernel"
    assert bench.crt_kernel_name(804) == "crt_product_kernel"


@pytest.mark.parametrize("flag", ["--water", "--graph-precision"])
def test_graph_workload_flags_are_declared(flag):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=300, cwd=str(ROOT))
    assert r.returncode == 0 and flag in r.stdout
