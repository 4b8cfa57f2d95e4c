"""The C++ BlockManager facade (include/bcm) driving the GPU: a sor1-style
driver reproduces examples/sor1/log.4:54 in host- and device-resident mode;
a driver with its own Divider and boundary setter exchanges a Vector3D
field (each component as the Scalar3D exchange) and reproduces updateVC
face by face through VCUpdater::copyFromNeighbor."""
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


def test_vector_field_and_per_face_exchange_on_facade(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "facade_vector")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    print("[gpu case] tests/cpp/facade_vector", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    print(out.stdout[-3000:])
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    assert "PASS" in out.stdout
    assert (tmp_path / "layout.txt").read_text().startswith("Block Layout Information")
