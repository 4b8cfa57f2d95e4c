"""bcm_jacobi_poisson_iterate: nsweeps of (jacobi7_poisson, ghost exchange)
in one call equal the same steps made one call at a time, bit for bit on
every cell -- the one-launch LDS-resident form (config 1's 64 x 16^3, an
adaptive 456 x 8^3 mesh with level changes) and the halo-gathering form on
compact iterates for meshes too large to be resident (2,528 x 16^3 and
32^3 blocks without boundary conditions, whose stored random ghosts enter
the unfolded stencil; shell meshes whose outer faces are Dirichlet, folded
into the coefficients; two and three steps, where the first and the last
step are the only halo steps or the only middle one)."""
import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh, Tree
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def stepwise(mesh, x, b, omega, nsweeps):
    a = mesh.field(np.float64, x)
    c = mesh.field(np.float64, x)
    for _ in range(nsweeps):
        mesh.jacobi7_poisson(c, a, b, omega)
        a, c = c, a
        mesh.update_vc(a)
    return a.download()


@pytest.mark.parametrize("name,n,nsweeps", [("config1", 16, 100), ("config1", 16, 37), ("simple23", 8, 25),
                                            ("simple24", 16, 6), ("simple12_n32", 32, 5), ("simple12_n32", 32, 2),
                                            ("shell5", 16, 7), ("shell3", 32, 3)])
def test_iterate_equals_stepwise_calls(name, n, nsweeps):
    if name == "config1":
        blocks = W.config1_tree().blocks
    elif name == "simple23":
        blocks = Tree(divider=L.DIV_SIMPLE, min_level=2, max_level=3).blocks
    elif name == "simple24":
        blocks = Tree(divider=L.DIV_SIMPLE, min_level=2, max_level=4).blocks
    elif name == "shell5":
        blocks = W.shell_tree(5).blocks
    elif name == "shell3":
        blocks = W.shell_tree(3).blocks
    else:
        blocks = Tree(divider=L.DIV_SIMPLE, min_level=1, max_level=2).blocks
    case(f"jacobi_poisson_iterate {name} nb={len(blocks)} n={n} sweeps={nsweeps}")
    vc = 1
    mesh = Mesh(blocks, n, vc)
    rng = np.random.default_rng(n + nsweeps)
    w = n + 2 * vc
    x = rng.standard_normal((len(blocks), w, w, w))
    rhs = mesh.field(np.float64, W.rhs_field(blocks, n, vc, seed=2))
    b = mesh.field(np.float64)
    mesh.poisson_assemble(None, b, rhs, [0.25])
    want = stepwise(mesh, x, b, 0.9, nsweeps)
    f = mesh.field(np.float64, x)
    mesh.jacobi_poisson_iterate(f, b, 0.9, nsweeps)
    got = f.download()
    np.testing.assert_array_equal(got, want)
    # The first step against the oracle (jacobi_smoother with the folded
    # Poisson coefficients, then the ghost exchange).
    if nsweeps == 25:
        A = [np.zeros_like(x) for _ in range(7)]
        bo = np.zeros_like(x)
        oracle.poisson_assemble(blocks, n, vc, A, bo, W.rhs_field(blocks, n, vc, seed=2), [0.25])
        one = mesh.field(np.float64, x)
        mesh.jacobi_poisson_iterate(one, b, 0.9, 1)
        x1 = x.copy()
        oracle.jacobi7(len(blocks), n, vc, x1, x, A, bo, 0.9)
        oracle.ghost_fill(blocks, n, vc, x1)
        np.testing.assert_array_equal(one.download(), x1)
    mesh.close()


def test_iterate_zero_sweeps_and_bad_arguments():
    blocks = W.config1_tree().blocks
    mesh = Mesh(blocks, 16, 1)
    x = mesh.field(np.float64, np.ones((len(blocks), 18, 18, 18)))
    b = mesh.field(np.float64)
    mesh.jacobi_poisson_iterate(x, b, 1.0, 0)
    assert np.all(x.download() == 1.0)
    assert mesh.lib.bcm_jacobi_poisson_iterate(x.handle, b.handle, 1.0, -1) == L.ERR_ARG
    x32 = mesh.field(np.float32)
    assert mesh.lib.bcm_jacobi_poisson_iterate(x32.handle, b.handle, 1.0, 1) == L.ERR_ARG
    mesh.close()
