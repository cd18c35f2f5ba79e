"""bench.py's command line without a GPU: more than one GPU is refused with
a message and exit status 2, unknown path workloads are rejected before any
work, and the workload table names what the help text lists."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BENCH = os.path.join(HERE, "..", "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_unknown_path_workload_is_rejected():
    r = _run("--paths", "no_such_workload")
    assert r.returncode != 0 and "no_such_workload" in r.stderr


def test_help_lists_the_path_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for name in ("box256_pt", "showcase1080_pt_nee", "showcase1080_bdpt", "showcase1080_bdpt_mis",
                 "showcase1080_bdpt_lt"):
        assert name in r.stdout
