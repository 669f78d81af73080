"""Host-side checks of bench.py: the one-GPU refusal, the CPU-baseline record,
and the byte accounting behind ``roofline`` (no GPU needed)."""
import json
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_record_at_a_small_size():
    a = types.SimpleNamespace(cpu_sample_nz=8, nz=8, nx=32, re_tau=395.0, cpu_cores=2)
    cb = bench.cpu_baseline(a)
    assert set(cb) == {"value", "unit", "cores", "kind", "sample"}
    assert cb["kind"] == "port" and cb["cores"] == 2 and cb["value"] > 0
    assert cb["unit"] == "grid-pt RK3 steps/s" and "32x32x8" in cb["sample"]


def test_algorithmic_bytes_of_the_stage_kernels():
    n, s = 256 * 256 * 128, 4
    per = {k: bench.kernel_bytes(k, n, s) / (n * s)
           for k in ("k_col_synth", "k_divz_band", "k_rows", "k_col_anal", "k_press", "k_erows",
                     "k_sgs_cols_in", "k_sgs_rows", "k_sgs_cols_out")}
    assert per["k_sgs_rows"] == 13 and per["k_rows"] == 24 and per["k_col_anal"] == 16
    assert bench.kernel_bytes("k_rows", n, s, lean=True) / (n * s) == 19.5
    assert bench.kernel_bytes("k_wall", n, s) is None
    # the counter-measured HBM traffic of the profiled kernels is at least
    # their algorithmic bytes and within 6 % of them
    for (k, nx, nz, _), traffic in bench.MEASURED_TRAFFIC.items():
        name = "k_rows" if k in ("k_rows_mlp", "k_rows_lean") else k
        alg = bench.kernel_bytes(name, nx * nx * nz, 4, lean=k == "k_rows_lean")
        assert alg <= traffic <= 1.06 * alg, (k, traffic, alg)


@pytest.mark.parametrize("kernel", sorted({k for (k, _, _, _) in bench.MEASURED_VALU}))
def test_valu_roofline_fraction_is_a_fraction(kernel):
    a = types.SimpleNamespace(nx=256, nz=128)
    v = bench.valu_roofline(kernel, a, 1.0)
    assert v["peak"] == 256 * 4 * 2.4e9 / 4
    assert v["achieved"] == pytest.approx(v["insts_per_launch"] / 1e-3)
