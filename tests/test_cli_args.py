"""Argument handling of the `gammaray` driver (no GPU: it exits before
creating a context)."""
import os
import subprocess

import pytest

CLI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "gammaray_amd", "gammaray")


@pytest.mark.parametrize("args,rc", [(["--help"], 0), (["--integrator", "nope"], 2), (["--bogus"], 2),
                                     (["--height", "0"], 2), (["--iterations"], 2)])
def test_arguments(args, rc):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == rc
    assert "usage: gammaray" in r.stderr
