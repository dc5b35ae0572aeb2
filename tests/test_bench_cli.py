"""bench.py's command line on the CPU: more than one GPU is refused with an
explanation before anything touches a device (the path shards by request and
has no collective), and the options the measurement relies on parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""


def test_defaults():
    sys.path.insert(0, ROOT)
    try:
        import bench
        saved = sys.argv
        sys.argv = ["bench.py"]
        try:
            args = bench.parse()
        finally:
            sys.argv = saved
    finally:
        sys.path.remove(ROOT)
    assert args.gpus == 1 and args.batch == 2048 and args.two_streams
    assert [int(v) for v in args.sweep.split(",")] == [512, 1024, 4096]
