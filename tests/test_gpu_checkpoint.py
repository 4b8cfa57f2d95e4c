"""Checkpointed serving.

bcm_server_checkpoint writes a running request's current iterate in
SolverTest's Dump2 restart format (examples/SolverTest/LocalScalar3D.cpp:
227-281) without changing its solve; bcm_server_submit_restart resumes a
request from such a file on another server.

What pins the checkpoint's content: the engine's iterate after k
iterations is x + dx (the FP64 correction not yet folded in), and a solve
that stops at its iteration limit k folds dx into x in its last iteration
and gathers the ghosts, so the checkpoint taken after k iterations must
equal, cell for cell, the solution of the same request run with
max_iter = k, and the oracle's residual of the checkpoint must be the
residual that solve reports.
"""
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Server
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W
from oracle import restart as R

pytestmark = pytest.mark.gpu

N = 16  # block edge of the resume test


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_relres(blocks, x, rhs, n=N):
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, [0.0])
    r = np.zeros_like(rhs)
    oracle.calc_r(len(blocks), n, 1, r, A, x, b)
    ri, bi = W.interior(r, 1), W.interior(b, 1)
    return float(np.sqrt(np.sum(ri * ri)) / np.sqrt(np.sum(bi * bi)))


def drive(sv, tags):
    out = {}
    while len(out) < len(tags):
        sv.progress(0.05)
        for r in sv.poll():
            out[r["tag"]] = r
    return out


def run(reqs, max_iter=20000, batch_blocks=0, restart=None, n=N, precision="fp64"):
    sv = Server(64e9, max_active=4, n=n, check_every=4, max_iter=max_iter, batch_blocks=batch_blocks)
    for tag, blocks, rhs in reqs:
        sv.submit(blocks, 1e-10, precision, rhs=rhs, tag=tag, want_x=True, restart=restart)
    out = drive(sv, [r[0] for r in reqs])
    sv.close()
    return out


def read_ckpt(path, nblocks, n=N):
    return R.load2(np.zeros((nblocks, n + 2, n + 2, n + 2)), n, 1, path)


def requests(n=N):
    out = []
    for tag, seed, lv in ((1, 21, 3), (2, 24, 2)):
        blocks = W.random_tree(seed, max_level=lv).blocks
        out.append((tag, blocks, -W.rhs_field(blocks, n, 1, seed=70 + tag)))
    return out


@pytest.mark.parametrize("precision,n", [("fp64", 16), ("mixed", 16), ("fp64", 32)])
def test_checkpoints_of_a_batched_group_are_the_iterates(tmp_path, precision, n):
    reqs = requests(n)
    case(f"checkpoint: group of {[len(r[1]) for r in reqs]} blocks of {n}^3, {precision}, "
         "checkpoints after 0/4/8/12 iterations")
    plain = run(reqs, n=n, precision=precision)
    d = str(tmp_path / "BIN")
    # A small request holds the only slot while the two others queue, so
    # they are admitted together as one batched group when it finishes.
    sv = Server(64e9, max_active=1, n=n, check_every=4, batch_blocks=4096)
    blocker = W.random_tree(23, max_level=2).blocks
    sv.submit(blocker, 1e-6, "fp64", seed=5, tag=0)
    for tag, blocks, rhs in reqs:
        sv.submit(blocks, 1e-10, precision, rhs=rhs, tag=tag, want_x=True)
    first = []
    while not first:
        sv.progress(0.0)  # one pass: at most one more block of iterations
        first = sv.poll()
    assert first[0]["tag"] == 0 and sv.stats()["queued"] == 0 and sv.stats()["active"] == 2
    its = {1: [], 2: []}
    for k in range(4):
        for tag in (1, 2):
            its[tag].append(sv.checkpoint(tag, d, f"x{tag}", k))
        sv.progress(0.0)  # the checkpoint waited for the group: one more block of 4 iterations
    assert its[1] == its[2] == [0, 4, 8, 12]
    done = drive(sv, [1, 2])
    sv.close()
    # Checkpoints read the solve and change nothing in it.
    assert done[1]["t_admit"] == done[2]["t_admit"]
    for tag in (1, 2):
        assert done[tag]["status"] == plain[tag]["status"] == 1
        assert done[tag]["iters"] == plain[tag]["iters"]
        np.testing.assert_array_equal(done[tag]["x"], plain[tag]["x"])
    # The checkpoint after 8 iterations is the iterate of a solve stopped there.
    stopped = run(reqs, max_iter=8, n=n, precision=precision)
    for tag, blocks, rhs in reqs:
        ck = read_ckpt(R.name2(d, f"x{tag}", 0, 2), len(blocks), n)
        assert stopped[tag]["iters"] == 8
        np.testing.assert_array_equal(ck, stopped[tag]["x"])
        # The stopped solve's true residual, recomputed by the oracle from the
        # checkpoint (the residual of Jacobi-BiCGSTAB first rises above 1).
        rel = oracle_relres(blocks, ck, rhs, n)
        assert abs(rel - stopped[tag]["relres"]) <= 1e-9 * rel, (rel, stopped[tag]["relres"])
        # Before any iteration the iterate is the zero initial guess.
        ck0 = read_ckpt(R.name2(d, f"x{tag}", 0, 0), len(blocks), n)
        assert not np.any(ck0)


def test_resume_from_a_checkpoint_on_a_new_server(tmp_path):
    reqs = requests()[:1]
    tag, blocks, rhs = reqs[0]
    case(f"resume: {len(blocks)} blocks, FP64, tol 1e-10")
    plain = run(reqs)[tag]
    d = str(tmp_path / "BIN")
    sv = Server(64e9, max_active=1, n=N, check_every=4)
    sv.submit(blocks, 1e-10, "fp64", rhs=rhs, tag=tag)
    # A second request waits for the slot: it has no iterate yet.
    sv.submit(blocks, 1e-10, "fp64", rhs=rhs, tag=99)
    with pytest.raises(L.BCMError) as e:
        sv.checkpoint(99, d, "q", 0)
    assert e.value.code == L.ERR_STATE
    with pytest.raises(L.BCMError) as e:
        sv.checkpoint(12345, d, "q", 0)
    assert e.value.code == L.ERR_ARG
    assert sv.checkpoint(tag, d, "x", 0) == 0
    for _ in range(5):
        sv.progress(0.0)  # launches the next block of 4 iterations
        it = sv.checkpoint(tag, d, "x", 1)  # waits for it
    assert it == 20
    sv.close()  # the server goes away with both requests unfinished
    # Resumed from the zero iterate: the same solve, bit for bit.
    again = run(reqs, restart=(d, "x", 0))[tag]
    assert again["iters"] == plain["iters"]
    np.testing.assert_array_equal(again["x"], plain["x"])
    # Resumed after 20 iterations: converges to the tolerance from there.
    res = run(reqs, restart=(d, "x", 1))[tag]
    assert res["status"] == 1
    assert oracle_relres(blocks, res["x"], rhs) <= 1e-10 * (1 + 1e-6)
    # A restart file of another block count is refused at submission.
    sv = Server(64e9, max_active=1, n=N)
    other = requests()[1][1]
    with pytest.raises(L.BCMError) as e:
        sv.submit(other, 1e-10, "fp64", seed=1, tag=3, restart=(d, "x", 1))
    assert e.value.code == L.ERR_ARG
    sv.close()
