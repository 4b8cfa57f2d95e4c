"""Request server (continuous batching): every result equals the same
request solved alone, bit for bit, and meets its tolerance on the true FP64
residual recomputed by the oracle; admission respects the HBM budget."""
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Server
from bcmtools_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_relres(blocks, n, x, rhs):
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, [0.0])
    r = np.zeros_like(rhs)
    oracle.calc_r(len(blocks), n, 1, r, A, x, b)
    ri, bi = W.interior(r, 1), W.interior(b, 1)
    return float(np.sqrt(np.sum(ri * ri)) / np.sqrt(np.sum(bi * bi)))


def _requests():
    reqs = []
    for i, (seed, lv, tol, prec) in enumerate(((7, 3, 1e-10, "mixed"), (8, 4, 1e-6, "fp32"), (9, 2, 1e-10, "fp64"),
                                               (10, 3, 1e-10, "mixed"), (11, 4, 1e-10, "mixed"),
                                               (12, 2, 1e-6, "fp32"))):
        blocks = W.random_tree(seed, max_level=lv).blocks
        reqs.append((i, blocks, tol, prec, -W.rhs_field(blocks, 16, 1, seed=40 + i)))
    return reqs


def _run(reqs, max_active, budget=64e9, batch_blocks=0, window_blocks=0, max_wait=0.0):
    sv = Server(budget, max_active=max_active, n=16, check_every=4, batch_blocks=batch_blocks,
                window_blocks=window_blocks, max_wait=max_wait)
    for tag, blocks, tol, prec, rhs in reqs:
        sv.submit(blocks, tol, prec, rhs=rhs, tag=tag, want_x=True)
    out = {}
    while len(out) < len(reqs):
        sv.progress(0.05)
        for r in sv.poll():
            out[r["tag"]] = r
    assert sv.stats()["bytes_in_use"] == 0.0
    sv.close()
    return out


@pytest.mark.parametrize("window_blocks,max_wait", [(0, 0.0), (64, 0.0), (64, 0.001)])
def test_server_results_same_alone_and_in_the_mix(window_blocks, max_wait):
    """Every launch order (each group as soon as it is ready; smallest
    groups first within a window of blocks in flight, with and without a
    wait limit) gives each request its result alone."""
    reqs = _requests()
    case(f"server mix of {len(reqs)} requests, blocks {[len(r[1]) for r in reqs]}, window {window_blocks}, "
         f"max_wait {max_wait}")
    mixed = _run(reqs, max_active=6, window_blocks=window_blocks, max_wait=max_wait)
    for req in reqs:
        tag, blocks, tol, prec, rhs = req
        alone = _run([req], max_active=1)[tag]
        m = mixed[tag]
        assert m["status"] == alone["status"] == 1, (m, alone)
        assert m["iters"] == alone["iters"] and m["relres"] == alone["relres"]
        np.testing.assert_array_equal(m["x"], alone["x"])
        assert oracle_relres(blocks, 16, m["x"], rhs) <= tol * (1 + 1e-6)


def test_batched_groups_give_each_request_its_result_alone():
    """batch_blocks merges small same-precision requests into one batched
    mesh; each request's solution, iterations and residual equal those of
    the request solved alone; a request above batch_blocks runs alone; the
    synthetic right-hand side (seeded on the device) is the same in a group."""
    reqs = _requests()
    # The three mixed-precision requests fit one group; the FP32 request
    # above the cap runs alone.
    cap = sum(len(r[1]) for r in reqs if r[3] == "mixed")
    assert cap < max(len(r[1]) for r in reqs if r[3] == "fp32")
    case(f"server batched groups cap={cap}, blocks {[len(r[1]) for r in reqs]}")
    grouped = _run(reqs, max_active=2, batch_blocks=cap)
    # Members of one group are admitted and finished together.
    tags = sorted(grouped)
    assert any(grouped[a]["t_done"] == grouped[b]["t_done"] and grouped[a]["t_admit"] == grouped[b]["t_admit"]
               for i, a in enumerate(tags) for b in tags[i + 1:])
    for req in reqs:
        tag = req[0]
        alone = _run([req], max_active=1)[tag]
        g = grouped[tag]
        assert g["status"] == alone["status"] == 1, (g, alone)
        assert g["iters"] == alone["iters"] and g["relres"] == alone["relres"]
        np.testing.assert_array_equal(g["x"], alone["x"])
    # Seeded synthetic right-hand sides, three requests of one precision.
    seeded = []
    for i in range(3):
        blocks = W.random_tree(20 + i, max_level=3).blocks
        seeded.append((100 + i, blocks))
    sv = Server(64e9, max_active=1, n=16, check_every=4, batch_blocks=100000)
    for tag, blocks in seeded:
        sv.submit(blocks, 1e-9, "fp64", seed=tag, tag=tag, want_x=True)
    out = {}
    while len(out) < len(seeded):
        sv.progress(0.05)
        for r in sv.poll():
            out[r["tag"]] = r
    sv.close()
    for tag, blocks in seeded:
        one = Server(64e9, max_active=1, n=16, check_every=4)
        one.submit(blocks, 1e-9, "fp64", seed=tag, tag=tag, want_x=True)
        res = None
        while res is None:
            one.progress(0.05)
            got = one.poll()
            res = got[0] if got else None
        one.close()
        assert out[tag]["iters"] == res["iters"] and out[tag]["relres"] == res["relres"]
        np.testing.assert_array_equal(out[tag]["x"], res["x"])


def test_server_budget_admission_and_refusal():
    blocks = W.random_tree(3, max_level=3).blocks
    small = W.random_tree(4, max_level=2).blocks
    case(f"server budget: blocks {len(blocks)}, {len(small)}")
    # Room for one request of the larger mesh at a time.
    probe = Server(1e12, max_active=4, n=16)
    budget = 1.5 * (len(blocks) * 18 ** 3 * 8 * 2 + len(blocks) * 4096 * (2 * 8 + 6 * 4) +
                    len(blocks) * 18 ** 3 * (2 * 8 + 4))
    probe.close()
    sv = Server(budget, max_active=4, n=16, check_every=4)
    for tag in range(3):
        sv.submit(blocks, 1e-8, "mixed", seed=tag, tag=tag)
    sv.submit(W.random_tree(5, max_level=6).blocks, 1e-8, "mixed", seed=9, tag=99)  # larger than the budget
    st = sv.stats()
    assert st["active"] == 1 and st["queued"] == 2
    done = {}
    while len(done) < 4:
        sv.progress(0.05)
        for r in sv.poll():
            done[r["tag"]] = r
    assert done[99]["status"] == -8  # -BCM_ERR_BUSY
    for tag in range(3):
        assert done[tag]["status"] == 1 and done[tag]["relres"] <= 1e-8
    # Served one after another: each admitted when the previous one finished.
    order = sorted(range(3), key=lambda t: done[t]["t_admit"])
    for a, b in zip(order, order[1:]):
        assert done[b]["t_admit"] >= done[a]["t_done"] - 1e-6
    # The synthetic right-hand side depends on the seed only.
    sv.close()


def test_server_rejects_a_pending_tag():
    """A tag names one pending request (the server reads its rhs and writes
    its x through the caller's buffers): a second submit with the same tag
    is refused by the library and by the Python mirror, and the tag can be
    used again once the first request's result has been polled."""
    import ctypes

    from bcmtools_amd import _lib as L
    blocks = W.random_tree(3, max_level=2).blocks
    case(f"server duplicate tag: {len(blocks)} blocks")
    sv = Server(64e9, max_active=2, n=16, check_every=4)
    sv.submit(blocks, 1e-8, "fp64", seed=1, tag=5)
    with pytest.raises(L.BCMError):
        sv.submit(blocks, 1e-8, "fp64", seed=2, tag=5)
    b = np.ascontiguousarray(blocks, dtype=L.BLOCK_DTYPE)
    rc = sv.lib.bcm_server_submit(sv.handle, b.ctypes.data, len(b), None, 2, ctypes.c_double(1e-8), 0, None, 5)
    assert rc == L.ERR_ARG
    done = []
    while not done:
        sv.progress(0.05)
        done += sv.poll()
    assert done[0]["tag"] == 5 and done[0]["status"] == 1
    sv.submit(blocks, 1e-8, "fp64", seed=3, tag=5)
    done = []
    while not done:
        sv.progress(0.05)
        done += sv.poll()
    assert done[0]["tag"] == 5 and done[0]["status"] == 1
    sv.close()
