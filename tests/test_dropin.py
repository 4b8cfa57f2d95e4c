"""The reference's examples/sor1 driver, unchanged, on the facade.

examples/sor1/{main,Solver,Timing}.cpp are compiled from the reference tree
with include/bcm in place of the reference library's headers (its own
Config.h, Solver.h, BoundaryConditionSetter.h and Timing.h stay first on
the include path) and linked with libbcm_hip.so.  The driver computes its
SOR sweeps on the host arrays (getData()) and calls updateVC, which the
facade runs on the GPU.

CPU test (where the reference tree is present): it compiles, and up to its
first device call it prints the reference's block layout of the case of
examples/sor1/log.4 (456 blocks; 96 / 2,488 / 96 faces at level
difference -1 / 0 / +1, log.4 summed over its four ranks).
GPU test (the binary built by __graft_entry__.build()): the whole run
reproduces log.4:54, errorMax = 0.00353051, with the exchange at once and
separately per axis.
"""
import os
import re
import subprocess

import pytest

from conftest import ROOT

REF_SOR1 = "/root/reference/examples/sor1"
DROPIN = os.path.join(ROOT, "tests", "cpp", "_dropin", "sor1")

# The case of examples/sor1/log.4 (the keys examples/sor1/Config.h reads).
CONF = """\
minLevel = 0
maxLevel = 3
treeType = simple
ordering = {ordering}
type = x
nLoopOuter = 400
nLoopInner = 1
omega = 1.2
size = 2
vc = 1
separateVCUpdate = {separate}
output = out.dat
"""


def write_conf(tmp_path, ordering="random", separate="false"):
    p = tmp_path / "sor1.conf"
    p.write_text(CONF.format(ordering=ordering, separate=separate))
    return str(p)


@pytest.mark.skipif(not os.path.isdir(REF_SOR1), reason="needs the reference tree (container only)")
def test_reference_sor1_compiles_unchanged_and_lays_out_the_log4_mesh(tmp_path):
    exe = str(tmp_path / "sor1")
    srcs = [os.path.join(REF_SOR1, f) for f in ("main.cpp", "Solver.cpp", "Timing.cpp")]
    cmd = ["g++", "-O2", "-std=c++17", "-DTIMING", "-I" + REF_SOR1, "-I" + os.path.join(ROOT, "include", "bcm"),
           "-I" + os.path.join(ROOT, "include"), "-o", exe] + srcs + [
           "-L" + os.path.join(ROOT, "bcmtools_amd"), "-lbcm_hip",
           "-Wl,-rpath," + os.path.join(ROOT, "bcmtools_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe, write_conf(tmp_path)], capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path)).stdout
    for line in ("0: [0:455] #456", "L=2: 8", "L=3: 448", "total: 456", "dLevel=-1: 96", "dLevel= 0: 2488",
                 "dLevel=+1: 96"):
        assert line in out, (line, out)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN), reason="the drop-in sor1 binary is built where the reference tree is")
@pytest.mark.parametrize("separate,ordering", [("false", "random"), ("true", "Z"), ("false", "Hilbert")])
def test_reference_sor1_reproduces_log4_on_the_gpu(tmp_path, separate, ordering):
    print(f"[gpu case] reference sor1 driver, separateVCUpdate={separate}, ordering={ordering}", flush=True)
    r = subprocess.run([DROPIN, write_conf(tmp_path, ordering, separate)], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"errorMax = (\S+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert m.group(1) == "0.00353051", m.group(1)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN), reason="the drop-in sor1 binary is built where the reference tree is")
def test_reference_sor1_with_three_inner_sweeps_on_the_gpu(tmp_path):
    """nLoopInner = 3 on a level 1-3 mesh of 4^3 blocks, z Dirichlet
    (b0 = -0.5, b1 = 1.25): the errorMax the reference library itself
    printed for this configuration (tests/golden/sor1_inner3_z.npz)."""
    from conftest import load_golden
    g = load_golden("sor1_inner3_z")
    conf = tmp_path / "inner3.conf"
    conf.write_text("minLevel = 1\nmaxLevel = 3\ntreeType = simple\nordering = Hilbert\ntype = z\n"
                    "nLoopOuter = 12\nnLoopInner = 3\nomega = 1.3\nsize = 4\nvc = 1\nb0 = -0.5\nb1 = 1.25\n"
                    "output = out.dat\n")
    print("[gpu case] reference sor1 driver, nLoopInner=3", flush=True)
    r = subprocess.run([DROPIN, str(conf)], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"errorMax = (\S+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert float(m.group(1)) == pytest.approx(float(g["error_max"]), rel=1e-6), (m.group(1), g["error_max"])
