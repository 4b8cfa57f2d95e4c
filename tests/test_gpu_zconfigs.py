"""The solver workloads at their full sizes.

Config 3: the level-7 mid-surface shell (26,888 blocks of 16^3, 110 M
cells), mixed FP32/FP64 BiCGSTAB to relative residual 1e-10; the oracle
recomputes ||b - A x|| / ||b|| on the CPU, and the error against the FP64
solution is stated.

Config 4: 256 random meshes (490,011 blocks of 16^3, 2.0 G cells) solved as
one batch to 1e-10: every segment converges on its true FP64 residual,
every 8th mesh solved alone gives the same bits, the smallest meshes'
residuals are recomputed by the oracle.

Config 5: a seeded request stream of the served workload (block counts
log-uniform in [64, 50k], at least two requests of 10k blocks or more,
tolerance 1e-6 or 1e-10 with the precision chosen by tolerance: FP32 for
1e-6, mixed for 1e-10) through the continuous-batching server: every
result is bit-identical to the same request served alone and meets its
tolerance by the oracle's residual.

These run for minutes; each prints its phases.
"""
import sys
import time

import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh, Server
from bcmtools_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 16


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_relres(blocks, n, x, rhs):
    """||b - A x|| / ||b|| of the folded Poisson system (Dirichlet 0), by the oracle."""
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, [0.0])
    r = np.zeros_like(rhs)
    oracle.calc_r(len(blocks), n, 1, r, A, x, b)
    ri, bi = W.interior(r, 1), W.interior(b, 1)
    return float(np.sqrt(np.sum(ri * ri)) / np.sqrt(np.sum(bi * bi)))


def test_config3_mixed_bicgstab_level7():
    blocks = W.config3_tree().blocks
    assert len(blocks) == 26888
    case(f"config 3: {len(blocks)} x {N}^3, mixed BiCGSTAB to 1e-10")
    rhs = -W.rhs_field_fast(blocks, N, 1, seed=3)
    mesh = Mesh(blocks, N, 1)
    g = mesh.field(np.float64, rhs)
    x = mesh.field(np.float64)
    t0 = time.perf_counter()
    st = mesh.solve(x, g, [0.0], precision="mixed", tol=1e-10, max_iter=10000)[0]
    mesh.sync()
    case(f"config 3 mixed: {st} in {time.perf_counter() - t0:.2f} s")
    assert st["status"] == "converged" and st["relres"] <= 1e-10, st
    xm = x.download()
    rr = oracle_relres(blocks, N, xm, rhs)
    case(f"config 3 oracle relres {rr:.3e}")
    assert rr <= 1e-10 * (1 + 1e-6), rr
    # The FP64 engine's solution of the same system, for the stated error.
    x64 = mesh.field(np.float64)
    st64 = mesh.solve(x64, g, [0.0], precision="fp64", tol=1e-10, max_iter=10000)[0]
    assert st64["relres"] <= 1e-10, st64
    a, b = W.interior(xm, 1), W.interior(x64.download(), 1)
    err = float(np.abs(a - b).max() / np.abs(b).max())
    case(f"config 3: max |x_mixed - x_fp64| / max |x_fp64| = {err:.3e}")
    assert err <= 1e-9


def test_config4_batch_of_256_meshes():
    M = 256
    tables = [W.random_tree(1000 + i).blocks for i in range(M)]
    counts = [len(t) for t in tables]
    assert sum(counts) == 490011
    case(f"config 4: {M} meshes, {sum(counts)} blocks ({min(counts)}..{max(counts)}), mixed to 1e-10")
    mesh = Mesh.batch(tables, N)
    x, rhs = mesh.field(np.float64), mesh.field(np.float64)
    t0 = time.perf_counter()
    for s in range(M):
        first, cnt = mesh.segment(s)
        assert cnt == counts[s]
        rhs.upload_blocks(first, -W.rhs_field_fast(tables[s], N, 1, seed=5000 + s))
        if s % 64 == 63:
            case(f"config 4: right-hand sides {s + 1}/{M} uploaded ({time.perf_counter() - t0:.1f} s)")
    mesh.set_progress(20.0)
    t0 = time.perf_counter()
    stats = mesh.solve(x, rhs, [0.0], precision="mixed", tol=1e-10, max_iter=10000)
    mesh.sync()
    case(f"config 4: batch solved in {time.perf_counter() - t0:.1f} s, max iters "
         f"{max(e['iters'] for e in stats)}, max relres {max(e['relres'] for e in stats):.3e}")
    for s, e in enumerate(stats):
        assert e["status"] == "converged" and e["relres"] <= 1e-10, (s, e)
    # Every 8th mesh solved alone: the same bits.
    for s in range(0, M, 8):
        first, cnt = mesh.segment(s)
        xb = x.download_blocks(first, cnt)
        alone = Mesh(tables[s], N)
        r1 = -W.rhs_field_fast(tables[s], N, 1, seed=5000 + s)
        xa = alone.field(np.float64)
        sa = alone.solve(xa, alone.field(np.float64, r1), [0.0], precision="mixed", tol=1e-10, max_iter=10000)[0]
        assert sa == stats[s], (s, sa, stats[s])
        np.testing.assert_array_equal(xb, xa.download())
        alone.close()
    case("config 4: every 8th mesh alone is bit-identical")
    # The smallest meshes: the oracle's residual.
    for s in sorted(range(M), key=lambda i: counts[i])[:6]:
        first, cnt = mesh.segment(s)
        r1 = -W.rhs_field_fast(tables[s], N, 1, seed=5000 + s)
        rr = oracle_relres(tables[s], N, x.download_blocks(first, cnt), r1)
        assert rr <= 1e-10 * (1 + 1e-6), (s, counts[s], rr)


def _stream(seed=11, R=8):
    """Requests of the served workload: block counts log-uniform in [64, 50k]
    (two of them forced to >= 10k blocks), tolerance 1e-6 or 1e-10, precision
    by tolerance."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(64), np.log(50000), size=R)).astype(int)
    targets[:2] = np.maximum(targets[:2], [12000, 20000])
    tols = rng.choice([1e-6, 1e-10], size=R)
    tols[:2] = [1e-10, 1e-6]
    reqs = []
    for i in range(R):
        blocks = W.serving_tree(7000 + i, int(targets[i])).blocks
        prec = "fp32" if tols[i] == 1e-6 else "mixed"
        reqs.append((i, blocks, float(tols[i]), prec, -W.rhs_field_fast(blocks, N, 1, seed=9000 + i)))
    return reqs


def _serve(reqs, max_active, window_blocks=0, max_wait=0.0):
    sv = Server(200e9, max_active=max_active, n=N, check_every=8, window_blocks=window_blocks, max_wait=max_wait)
    for tag, blocks, tol, prec, rhs in reqs:
        sv.submit(blocks, tol, prec, rhs=rhs, tag=tag, want_x=True)
    out = {}
    while len(out) < len(reqs):
        sv.progress(0.2)
        for r in sv.poll():
            out[r["tag"]] = r
    assert sv.stats()["bytes_in_use"] == 0.0
    sv.close()
    return out


def test_config5_served_stream_same_alone_and_in_the_mix():
    reqs = _stream()
    sizes = [len(r[1]) for r in reqs]
    assert sum(s >= 10000 for s in sizes) >= 2
    case(f"config 5: {len(reqs)} requests, blocks {sizes}, tols {[r[2] for r in reqs]}")
    t0 = time.perf_counter()
    # Served as the benchmark serves: smallest groups first within 8,192
    # blocks in flight, no group held back more than 5 s.
    mix = _serve(reqs, max_active=16, window_blocks=8192, max_wait=5.0)
    case(f"config 5: served in {time.perf_counter() - t0:.1f} s")
    for req in reqs:
        tag, blocks, tol, prec, rhs = req
        alone = _serve([req], max_active=1)[tag]
        m = mix[tag]
        case(f"config 5 request {tag}: {len(blocks)} blocks, {prec}, tol {tol:g}: {m['iters']} its, "
             f"relres {m['relres']:.2e}")
        assert m["status"] == alone["status"] == 1, (tag, m["status"], alone["status"])
        assert m["iters"] == alone["iters"] and m["relres"] == alone["relres"]
        np.testing.assert_array_equal(m["x"], alone["x"])
        assert oracle_relres(blocks, N, m["x"], rhs) <= tol * (1 + 1e-6)
