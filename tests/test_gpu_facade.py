"""The C++ BlockManager facade (include/bcm) driving the GPU: a sor1-style
driver reproduces examples/sor1/log.4:54 in host- and device-resident mode."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_sor1_driver_on_facade():
    exe = os.path.join(ROOT, "tests", "cpp", "sor1_facade")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    print("[gpu case] tests/cpp/sor1_facade", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASS" in out.stdout
