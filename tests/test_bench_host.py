"""bench.py's host-side contract (no GPU needed)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    """Every box here has one GPU: --gpus above 1 exits non-zero with a
    message, before touching a device."""
    r = _bench("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_unknown_workload_is_refused():
    r = _bench("--workload", "no-such-workload", "--no-cpu")
    assert r.returncode != 0
    assert "unknown workload" in (r.stdout + r.stderr)


def test_traffic_figures_are_read_from_the_committed_counter_summary():
    """The roofline's traffic comes from the committed counter summary under
    profiles/, one line per kernel kind."""
    sys.path.insert(0, ROOT)
    import bench
    t_fwd = bench.measured_traffic("grid3d128", "fp64", "tri_fwd")
    t_bwd = bench.measured_traffic("grid3d128", "fp64", "tri_bwd")
    assert t_fwd and t_bwd and 5e7 < t_fwd < 2e8 and 5e7 < t_bwd < 2e8
    assert bench.measured_traffic("grid3d128", "fp64", "spmm") is None
    assert bench.measured_traffic("no-such-workload", "fp64", "tri_fwd") is None
