"""Host-side checks of bench.py's command line (no GPU needed): more than one GPU is refused with a message and
a non-zero status before anything touches a device."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120,
                       cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert not r.stdout.strip().startswith("{")  # no result line


def test_cpu_baseline_leg_on_the_host():
    """The bench's CPU-baseline leg (the oracle's C restatement, its single-thread run, the numpy port, the H^2 shape
    and the env update) runs without a GPU and reports the contract's fields, on a small bounded sample."""
    sys.path.insert(0, str(ROOT))
    import bench
    from oracle import heff_c as OC

    if not OC.host_supported():
        import pytest

        pytest.skip("the C oracle needs AVX2 and FMA")
    if not OC.LIB_PATH.exists():
        from oracle.build import build

        build()
    out = bench.cpu_baseline(chi=64, budget_s=0.2, max_reps=3)
    assert out["kind"] == "port" and out["unit"] == "TFLOP/s" and out["value"] > 0 and out["cores"] >= 1
    for key in ("single_thread", "numpy_blas", "h2_matvec", "env_update"):
        assert out[key]["value"] > 0, key
    assert out["host"]["nproc"] >= 1
