"""Host-side checks of bench.py that need no GPU: the algorithmic-bytes model against the
worked examples of the measurement plan (two vesicles at 1024^2 with 24 psi iterations,
two ellipsoids at 128^3), and the multi-GPU refusal."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_algorithmic_bytes_worked_examples(jv):
    import bench
    c3 = jv.workload_c3().conf
    c4 = jv.workload_c4().conf
    # worked examples: 1024^2, N = 2, k = 24: ~20.5 GB (FP32) and ~41 GB (FP64) per step;
    # 128^3, N = 2: ~5.9 GB (FP32) and ~11.8 GB (FP64) per step
    assert bench.algorithmic_bytes(c3, 4, 24) / 1e9 == pytest.approx(20.5, rel=0.01)
    assert bench.algorithmic_bytes(c3, 8, 24) / 1e9 == pytest.approx(41.0, rel=0.01)
    assert bench.algorithmic_bytes(c4, 4, 0) / 1e9 == pytest.approx(5.9, rel=0.01)
    assert bench.algorithmic_bytes(c4, 8, 0) / 1e9 == pytest.approx(11.8, rel=0.01)
    # each psi iteration adds 88 real N-fields
    d = bench.algorithmic_bytes(c3, 4, 25) - bench.algorithmic_bytes(c3, 4, 24)
    assert d == 88 * 4 * c3.Nx * c3.Ny * c3.N


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_built_kernel_byte_model(jv):
    import bench
    c3 = jv.workload_c3().conf
    u = 8 * c3.Nx * c3.Ny * c3.N
    # one more psi iteration adds 72 real N-fields in FP64 (76 in FP32) with the built kernels
    assert bench.algorithmic_bytes_built(c3, 8, 25) - bench.algorithmic_bytes_built(c3, 8, 24) == 72 * u
    assert bench.algorithmic_bytes_built(c3, 4, 25) - bench.algorithmic_bytes_built(c3, 4, 24) == 76 * u // 2
    # outside the iterations the built step moves fewer bytes than the survey model counts
    # (spectral inner products, no back-transform of the split solutions)
    assert bench.algorithmic_bytes_built(c3, 8, 0) == 225 * u + 82 * u // c3.N
    c4 = jv.workload_c4().conf
    u4 = 8 * c4.Nx * c4.Ny * c4.Nz * c4.N
    assert bench.algorithmic_bytes_built(c4, 8, 0) == 181 * u4 + 130 * u4 // c4.N
    assert bench.algorithmic_bytes_built(c4, 8, 0) < bench.algorithmic_bytes(c4, 8, 0)
