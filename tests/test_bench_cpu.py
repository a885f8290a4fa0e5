"""The benchmark's reporting logic on the host: the single-GPU contract, the
kernel-instance names matched against rocprofv3 traces, and the roofline
object's arithmetic."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("name,prec,variant,expect", [
    ("coly_gauge", 64, (3, 5, 5), "gpue::col_sweep<double, 512, 5, 3>"),
    ("coly_gauge+norm", 64, (3, 5, 5), "gpue::col_sweep<double, 512, 5, 3>"),
    ("coly_k", 32, (0, 2, 0), "gpue::col_sweep<float, 512, 2, 2>"),
    ("row_fwd+vpre", 64, (3, 5, 5), "gpue::row_sweep<double, 512, 3, 0>"),
    ("row_gauge_fwd+vpre", 64, (5, 1, 1), "gpue::row_sweep<double, 512, 3, 6>"),   # x: 4, 5 act as 2, 3
    ("row_inv+vpost+gx", 64, (1, 1, 1), "gpue::row_sweep<double, 512, 1, 1>"),
    # closing sweeps with no gauge operator: the instance compiled without it
    ("row_inv+vpost+norm", 32, (0, 0, 0), "gpue::row_sweep<float, 512, 0, 65>"),
    ("row_inv_fwd+vpost+vpre", 64, (3, 5, 5), "gpue::row_sweep<double, 512, 3, 69>"),
    ("row_gauge+norm", 64, (2, 0, 0), "gpue::row_sweep<double, 512, 2, 2>"),
])
def test_kernel_symbols(name, prec, variant, expect):
    assert bench.kernel_symbol(name, prec, (512, 512, 1), variant) == expect


def test_kernel_symbol_z_axis():
    assert bench.kernel_symbol("colz_k", 64, (256, 256, 256), (3, 1, 2)) == "gpue::col_sweep<double, 256, 2, 2>"


def test_roofline_object():
    rows = [{"kernel": "coly_gauge", "symbol": "gpue::col_sweep<double, 512, 5, 3>", "us_per_launch": 8.0}]
    r = bench.roofline(rows, 2 * 16 * 512 * 512, "512sq")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(2 * 16 * 512 * 512 / 8e-6 / 1e9, abs=0.05)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, abs=1e-4)
    assert r["traffic"] == bench.MEASURED_TRAFFIC["gpue::col_sweep<double, 512, 5, 3>@512sq"]
    # an instance without a counter run reports no traffic rather than a wrong one
    rows[0]["symbol"] = "gpue::col_sweep<double, 512, 4, 3>"
    assert bench.roofline(rows, 1, "512sq")["traffic"] is None
    assert bench.roofline([], 1, "512sq") is None
