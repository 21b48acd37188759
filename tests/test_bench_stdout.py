"""bench.py prints one JSON line on stdout; what native code prints to fd 1
inside `stdout_to_stderr()` (rtm_routine's progress banner) goes to stderr."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_native_prints_go_to_stderr():
    code = ("import ctypes, bench\n"
            "with bench.stdout_to_stderr():\n"
            "    ctypes.CDLL(None).printf(b'native banner\\n')\n"
            "print('result line')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=REPO, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == ["result line"]
    assert "native banner" in r.stderr


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=REPO, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "exactly one GPU" in r.stderr
