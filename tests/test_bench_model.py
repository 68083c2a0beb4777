"""The benchmark's byte model and its --gpus contract (no GPU work)."""
import importlib.util
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


# Compulsory bytes per forward + VJP as tabulated for the workloads (FP64, FP32), GB.
TABLE = {(64, 5): (0.394, 0.199), (256, 7): (6.32, 3.19), (1024, 9): (101.1, 51.05), (2048, 10): (404.4, 204.2)}


@pytest.mark.parametrize("nside, J", sorted(TABLE))
def test_compulsory_bytes_match_the_workload_table(bench, nside, J):
    f64, f32 = TABLE[(nside, J)]
    assert bench.compulsory_bytes(nside, J, 8) / 1e9 == pytest.approx(f64, rel=2e-3)
    assert bench.compulsory_bytes(nside, J, 4) / 1e9 == pytest.approx(f32, rel=2e-3)


def test_eval_roofline_fractions(bench):
    r = bench.eval_roofline(1024, 9, 4, 100.0, 5000.0)
    assert r["achieved_GBps"] == pytest.approx(5105.1, rel=1e-4)
    assert r["frac_of_peak"] == pytest.approx(5105.1 / 8000.0, rel=1e-3)
    assert r["frac_of_measured_copy"] == pytest.approx(5105.1 / 5000.0, rel=1e-3)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr
