"""bench.py's GPU-free modes run on the CPU: its imports and helpers stay
sound between GPU runs."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_counter_summary(tmp_path):
    csv = tmp_path / "c.csv"
    csv.write_text("Kernel_Name,Counter_Name,Counter_Value\n"
                   '"rows_kernel<64, false>",FETCH_SIZE,10\n"rows_kernel<64, false>",FETCH_SIZE,30\n')
    r = run_bench("--summarize-counters", str(csv))
    assert r.returncode == 0, r.stderr
    assert "mean=" in r.stdout and "20.0" in r.stdout


def test_help_lists_the_modes():
    r = run_bench("--help")
    assert r.returncode == 0
    for flag in ("--serving", "--isochrones", "--table-2ch", "--autotune"):
        assert flag in r.stdout
