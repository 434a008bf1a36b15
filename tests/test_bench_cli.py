"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the library runs on one GPU), before anything touches a device."""
import pathlib
import subprocess
import sys

REPO = pathlib.Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr or "single MI355X" in p.stderr
    assert p.stdout == ""
