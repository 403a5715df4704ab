"""bench.py's host-side contract (no GPU needed)."""
import os
import subprocess
import sys

import pytest

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
    t_fwd = bench.measured_traffic("grid3d128", "fp64", "tri_fwd", 48)
    t_bwd = bench.measured_traffic("grid3d128", "fp64", "tri_bwd", 48)
    assert t_fwd and t_bwd and 5e7 < t_fwd < 2e8 and 5e7 < t_bwd < 2e8
    assert bench.measured_traffic("grid3d128", "fp64", "spmm", 48) is None
    assert bench.measured_traffic("no-such-workload", "fp64", "tri_fwd", 48) is None
    # a pass of another width is not the one the counters measured
    assert bench.measured_traffic("grid3d128", "fp64", "tri_fwd", 192) is None
    # the c3 served pass with FP16 preconditioner rows: backward level launches
    # move about half a gigabyte; a pass with FP32 rows is not the one measured
    assert 3e8 < bench.measured_traffic("chunglu5m", "mixed", "tri_bwd", 256, rows="fp16") < 1e9
    assert bench.measured_traffic("chunglu5m", "mixed", "tri_bwd", 256, rows="fp32") is None


@pytest.mark.parametrize("xyz", [False, True])
def test_trace_summary_takes_the_window_between_the_markers(tmp_path, xyz):
    """--trace-summary keeps the kernels between the two marker copies
    (a warm and a timed 512-item k_copy_f64 on either side of the profiled
    pass) and nothing outside them; the 1 GiB copy is no marker.  The grid
    comes as one column or as X, Y, Z."""
    grid_cols = '"Grid_Size_X","Grid_Size_Y","Grid_Size_Z"' if xyz else '"Grid_Size"'
    head = ('"Correlation_Id","Dispatch_Id","Kernel_Name",%s,"Start_Timestamp","End_Timestamp"\n' % grid_cols)
    rows = [("ss::kb_sketch(x)", 1000, 0, 500),                 # before the window
            ("ss::k_copy_f64(long, double const*, double*)", 512, 1000, 1010),
            ("ss::k_copy_f64(long, double const*, double*)", 512, 1020, 1030),
            ("void ss::kb_spmm_dot<float>(a)", 4096, 2000, 4000),
            ("void ss::kb_spmm_dot<float>(a)", 4096, 5000, 6000),
            ("void ss::k_trisolve_level<float, float, float, true, 1>(b)", 64, 6000, 6500),
            ("ss::k_copy_f64(long, double const*, double*)", 512, 7000, 7010),
            ("ss::k_copy_f64(long, double const*, double*)", 512, 7020, 7030),
            ("ss::k_copy_f64(long, double const*, double*)", 1 << 27, 8000, 9000)]
    p = tmp_path / "run_kernel_trace.csv"
    gfmt = (lambda g: "%d,1,1" % g) if xyz else (lambda g: "%d" % g)
    p.write_text(head + "".join('%d,%d,"%s",%s,%d,%d\n' % (i, i, n, gfmt(g), a, b) for i, (n, g, a, b) in enumerate(rows)))
    r = _bench("--trace-summary", str(p))
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "| `ss::kb_spmm_dot` | 2 | 0.00 | 1.5 |" in out
    assert "| `ss::k_trisolve_level (bwd)` | 1 |" in out
    assert "kb_sketch" not in out and "k_copy_f64" not in out


def test_pmc_summary_kinds(tmp_path, capsys):
    """The counter summary calibrates on the probe's copy and sorts level
    dispatches into kinds: backward levels, forward levels, and the forward
    core head (the level launch just before a forward core gather)."""
    import bench
    hdr = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
    lv = "void ss::k_trisolve_level<float, float __attribute__((ext_vector_type(2))), " \
         "_Float16 __attribute__((ext_vector_type(2))), %s, 2>(int)"
    hg = "void ss::k_trisolve_huge<float, float, float, %s>(int)"
    rows = [(lv % "false", 10), (hg % "false", 5), (lv % "false", 7),
            ("void ss::k_core_gather<float, _Float16, false>(long)", 1),
            (lv % "true", 20), ("void ss::kb_spmm_dot<float, float, float>(int)", 100),
            ("void ss::k_copy_f64(long, double const*, double*)", 1 << 19)]
    for name, scale in (("f.csv", 1), ("w.csv", 2)):
        with open(tmp_path / name, "w") as f:
            f.write(hdr)
            for i, (k, v) in enumerate(rows):
                f.write('%d,"%s",X,%d\n' % (i + 1, k, v * scale))
    assert bench.pmc_summary(str(tmp_path / "f.csv"), str(tmp_path / "w.csv")) == 0
    out = capsys.readouterr().out
    # copy: fetch 512 KiB-units -> factor 2; write 2^20 -> factor 1; traffic = 2 f + 1 w = 4 v KiB
    assert "traffic_bytes_per_launch tri_bwd: %d" % (4 * 20 * 1024) in out
    assert "traffic_bytes_per_launch tri_fwd: %d" % (4 * 10 * 1024) in out
    assert "traffic_bytes_per_launch spmm: %d" % (4 * 100 * 1024) in out
