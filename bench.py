"""Benchmark: red-black SOR sweeps with ghost fill on the adaptive FP64 mesh.

Workload (default): the adaptive shell octree refined to level 5 --
3,816 leaf blocks of 32^3 cells (125 M interior cells), one ghost layer,
RootGrid(2,2,2) of length 0.5, 2:1 balanced, Dirichlet u = 0 on the
outer faces -- solving lap(u) = s with s = -(sin(pi x) sin(pi y) sin(pi z)
+ 0.01 N(0,1)) by the sor2 red-black SOR at omega = 1.6.

One step = one outer iteration of examples/sor2/Solver.cpp:77-133:
red-black SOR on every block, Dirichlet mirror, ghost fill (same-level
copy + F2C + C2F), Dirichlet mirror.  Inputs are resident in HBM before
the timed region.  value = interior cell updates per second (GCUPS).

roofline: the dominant kernel is the halo-gathering red-black sweep
(k_rbsor_halo<double,32,true,false>: one sweep reading iterate A, writing
iterate B, ghost frame built from the neighbours' interiors); its
algorithmic traffic is 24 B per interior cell (read u and s, write u,
FP64), timed with HIP events on the mesh stream inside the timed region
over its own launches (the K steps of the call minus the first, which runs
k_rbsor_fused on the stored ghosts, and the last, which writes the ghosted
field; those two are reported beside it).

The same line carries the rest of the metric on bounded samples (left out
with --quick): sor_to_tolerance (the sor2 iteration to relres 1e-8 on the
config-2 generator at level 3), batched (every 8th mesh of the 256-mesh
batch: batched mixed and FP64 against one mesh at a time, fused and plain
host-driven FP64) and serving (a 12-request config-5 stream through the
continuous-batching server: solves/s, p50/p99 latency, and the same
requests one at a time).

cpu_baseline: the oracle (C, OpenMP, -O3 -ffp-contract=off) runs the same
step on the same mesh for a bounded number of steps on the host cores.

Other workloads (--workload): jacobi (64 x 16^3, Jacobi7 + ghost fill),
bicgstab (level-7 shell, 26,888 x 16^3: mixed FP32/FP64 BiCGSTAB to 1e-10
against the FP64 paths), batched (256 random meshes solved as one batch,
with a per-mesh comparison on every 8th mesh), serving (Poisson-arrival
request stream through the continuous-batching server), copy (HBM
calibration).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
                       [--workload sor|jacobi|bicgstab|batched|serving|copy|restart]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for ln in fp:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_profiles(name):
    """HBM bytes measured with rocprofv3 --pmc (profiles/traffic.json), if recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fp:
            return json.load(fp).get(name)
    except (OSError, ValueError):
        return None


def bench_sor(args):
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    omega = 1.6
    t = W.config2_tree()
    blocks, n, vc = t.blocks, 32, 1
    ncell = len(blocks) * n ** 3
    print(f"[bench] sor: {len(blocks)} blocks x {n}^3 = {ncell} cells", file=sys.stderr, flush=True)
    f = W.rhs_field(blocks, n, vc, seed=0)
    s = -f
    del f
    mesh = Mesh(blocks, n, vc)
    gf = mesh.field(np.float64)
    gs = mesh.field(np.float64, s)
    vals = [0.0]
    mesh.dirichlet_mirror(gf, vals)
    mesh.sor_iterate(gf, gs, omega, vals, args.warmup)
    mesh.sync()
    mesh.reset_timing()
    mesh.set_timing(True)
    mesh.sync()
    t0 = time.perf_counter()
    mesh.sor_iterate(gf, gs, omega, vals, args.steps)
    mesh.sync()
    t1 = time.perf_counter()
    mesh.set_timing(False)
    elapsed = t1 - t0
    halo_ms, halo_n = mesh.kernel_time(L.K_SOR_HALO)
    end_ms, end_n = mesh.kernel_time(L.K_SOR)
    gh_ms, gh_n = mesh.kernel_time(L.K_GHOST)
    bc_ms, bc_n = mesh.kernel_time(L.K_DIRICHLET)
    if halo_n == 0:
        raise RuntimeError("bench: the halo-gathering sweep did not run (steps < 3?)")
    sor_avg = halo_ms / halo_n * 1e-3
    algo_bytes = 24.0 * ncell
    achieved = algo_bytes / sor_avg / 1e9
    gcups = ncell * args.steps / elapsed / 1e9
    print(f"[bench] halo sweep avg {sor_avg * 1e3:.3f} ms ({achieved:.0f} GB/s) over {halo_n}, first+last sweep avg "
          f"{end_ms / max(end_n, 1):.3f} ms, ghost avg {gh_ms / max(gh_n, 1):.3f} ms, dirichlet avg "
          f"{bc_ms / max(bc_n, 1):.3f} ms", file=sys.stderr, flush=True)
    result = {
        "metric": "stencil GCUPS (FP64 red-black SOR + ghost fill, adaptive octree)",
        "value": gcups,
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "adaptive shell octree level 5, 3816 x 32^3 blocks, vc=1, sor2 RB-SOR omega=1.6 "
                               "+ Dirichlet + ghost fill per step", "blocks": len(blocks), "block_cells": n,
                   "interior_cells": ncell, "omega": omega},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles("rbsor_halo_n32"),
                     "kernel": "k_rbsor_halo<double,32,true,false> (HIP-event average over its launches: the "
                               "sweeps between a call's first and last)",
                     "avg_ms": sor_avg * 1e3, "launches": halo_n,
                     "algorithmic_bytes_per_launch": algo_bytes},
        "kernels_ms_per_step": {"rbsor_halo": halo_ms / max(args.steps, 1),
                                "rbsor_first_and_last": end_ms / max(args.steps, 1),
                                "ghost_fill": gh_ms / max(args.steps, 1), "dirichlet": bc_ms / max(args.steps, 1)},
        "first_and_last_sweep_avg_ms": end_ms / max(end_n, 1),
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_sor(blocks, n, vc, s, omega, vals, args.cpu_steps)
    return result


def sor_to_tolerance(level=3, n=32, omega=1.4, tol=1e-8, chunk=500, cap=400000):
    """The sor2 outer iteration run to relres ||s - lap_h f|| / ||s|| <= tol
    on the config-2 generator at a reduced level, checked every `chunk`
    sweeps (bcm_sor_residual_noalloc)."""
    import ctypes

    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    blocks = W.shell_tree(level).blocks
    mesh = Mesh(blocks, n, 1)
    sf = mesh.field(np.float64, -W.rhs_field_fast(blocks, n, 1, seed=0))
    f = mesh.field(np.float64)
    mesh.dirichlet_mirror(f, [0.0])
    r = ctypes.c_double()
    sweeps, rr = 0, float("inf")
    mesh.sync()
    t0 = time.perf_counter()
    while sweeps < cap:
        mesh.sor_iterate(f, sf, omega, [0.0], chunk)
        sweeps += chunk
        if mesh.lib.bcm_sor_residual_noalloc(f.handle, sf.handle, ctypes.byref(r)) != 0:
            raise RuntimeError("bcm_sor_residual_noalloc failed")
        rr = r.value
        if not np.isfinite(rr) or rr <= tol:
            break
    el = time.perf_counter() - t0
    ncell = len(blocks) * n ** 3
    print(f"[bench] sor to tolerance: {sweeps} sweeps, relres {rr:.3e}, {el:.2f} s", file=sys.stderr, flush=True)
    mesh.close()
    return {"workload": f"shell octree level {level} (the config-2 generator one level coarser per step down from "
                        f"5), {len(blocks)} x {n}^3 blocks, sor2 RB-SOR + Dirichlet + ghost fill per sweep",
            "blocks": len(blocks), "omega": omega, "tol": tol, "sweeps": sweeps, "relres": rr,
            "converged": bool(rr <= tol), "time_s": el, "GCUPS": ncell * sweeps / el / 1e9,
            "omega_1_6": "diverges: with ghosts frozen between exchanges, block-wise RB-SOR is unstable at "
                         "omega >= 1.5 on these meshes (relres 1e23 after 500 sweeps at 1.5, 1e47 at 1.6 on the "
                         "level-4 mesh); the compiled reference sor1 gives errorMax 5.8e31 at omega 1.6 and "
                         "0.334 at 1.4 on 64 x 16^3 after 400 sweeps",
            "full_size_note": "on the level-5 mesh itself (3,816 x 32^3) omega 1.4 reaches relres 9.4e-9 after "
                              "475,000 sweeps in 272 s (bench.py --workload sortol), too long for this line"}


def cpu_sor(blocks, n, vc, s, omega, vals, steps):
    import oracle

    ncell = len(blocks) * n ** 3
    f = np.zeros_like(s)
    oracle.dirichlet(blocks, n, vc, f, vals)
    oracle.sor_iterate(blocks, n, vc, f, s, omega, vals, 1)  # warm caches / threads
    t0 = time.perf_counter()
    oracle.sor_iterate(blocks, n, vc, f, s, omega, vals, steps)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": ncell * steps / dt / 1e9, "unit": "GCUPS", "cores": cores, "kind": "port",
            "sample": f"same mesh and step, {steps} steps ({dt:.1f} s) on {cores} threads of {cpu_model()}"}


def bench_jacobi(args):
    """Config 1: 64 x 16^3, 100-sweep Jacobi with ghost fill (L2-resident)."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    t = W.config1_tree()
    blocks, n, vc = t.blocks, 16, 1
    ncell = len(blocks) * n ** 3
    f = W.rhs_field(blocks, n, vc, seed=1)
    mesh = Mesh(blocks, n, vc)
    rhs = mesh.field(np.float64, -f)
    gb = mesh.field(np.float64)
    mesh.poisson_assemble(None, gb, rhs, [0.0])
    gx = mesh.field(np.float64)
    # Every step (one jacobi_smoother sweep, then the ghost exchange) of the
    # run in one call: the 64 blocks stay in LDS for the whole run.
    mesh.jacobi_poisson_iterate(gx, gb, 1.0, max(args.warmup, 1))
    mesh.sync()
    mesh.reset_timing()
    mesh.set_timing(True)
    t0 = time.perf_counter()
    mesh.jacobi_poisson_iterate(gx, gb, 1.0, args.steps)
    mesh.sync()
    el = time.perf_counter() - t0
    jms, jn = mesh.kernel_time(L.K_JACOBI)
    javg = jms / max(args.steps, 1) * 1e-3  # per step, from the call's HIP events
    achieved = 24.0 * ncell / javg / 1e9
    res = {"metric": "stencil GCUPS (FP64 Jacobi7 + ghost fill, config 1)", "value": ncell * args.steps / el / 1e9,
           "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "uniform level-2 octree, 64 x 16^3, matrix-free Jacobi7 + ghost fill, "
                                  "all steps in one bcm_jacobi_poisson_iterate call (blocks resident in LDS)"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "note": "algorithmic 24 B per cell update; the blocks live in LDS and the exchange in L2, "
                                "so the step is bound by the grid barrier, not by HBM"}}
    mesh.close()
    res["level7"] = jacobi_level7(args)
    if not args.no_cpu:
        import oracle

        A = [np.zeros_like(f) for _ in range(7)]
        b = np.zeros_like(f)
        oracle.poisson_assemble(blocks, n, vc, A, b, -f, [0.0])
        x, x0 = np.zeros_like(f), np.zeros_like(f)
        sweeps = 100
        t0 = time.perf_counter()
        for _ in range(sweeps):
            oracle.jacobi7(len(blocks), n, vc, x0, x, A, b, 1.0)
            x, x0 = x0, x
            oracle.ghost_fill(blocks, n, vc, x)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        res["cpu_baseline"] = {"value": ncell * sweeps / dt / 1e9, "unit": "GCUPS", "cores": cores, "kind": "port",
                               "sample": f"same mesh, {sweeps} sweeps of jacobi_smoother (stored A) + ghost fill "
                                         f"({dt:.2f} s) on {cores} threads of {cpu_model()}"}
    return res


def jacobi_level7(args, steps=20):
    """The same iteration at HBM scale: the level-7 shell mesh (26,888 x
    16^3, Dirichlet outer faces), too large to stay in LDS, so every step
    streams the compact iterate (halo-gathering schedule).  Also times the
    same steps made as one jacobi7_poisson and one ghost fill call each."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    t = W.config3_tree()
    blocks, n = t.blocks, 16
    ncell = len(blocks) * n ** 3
    mesh = Mesh(blocks, n, 1)
    rhs = mesh.field(np.float64, -W.rhs_field_fast(blocks, n, 1, seed=1))
    gb = mesh.field(np.float64)
    mesh.poisson_assemble(None, gb, rhs, [0.0])
    gx = mesh.field(np.float64)
    mesh.jacobi_poisson_iterate(gx, gb, 1.0, 3)
    mesh.sync()
    mesh.reset_timing()
    mesh.set_timing(True)
    t0 = time.perf_counter()
    mesh.jacobi_poisson_iterate(gx, gb, 1.0, steps)
    mesh.sync()
    el = time.perf_counter() - t0
    hms, hn = mesh.kernel_time(L.K_JACOBI_HALO)
    ems, en = mesh.kernel_time(L.K_JACOBI)
    gms, gn = mesh.kernel_time(L.K_GHOST)
    mesh.set_timing(False)
    havg = hms / max(hn, 1)
    achieved = 24.0 * ncell / (havg * 1e-3) / 1e9
    # The reference's step as two calls (sweep, then exchange), for comparison.
    gy = mesh.field(np.float64)
    k = 5
    mesh.sync()
    t1 = time.perf_counter()
    for _ in range(k):
        mesh.jacobi7_poisson(gy, gx, gb, 1.0)
        gx, gy = gy, gx
        mesh.update_vc(gx)
    mesh.sync()
    el2 = time.perf_counter() - t1
    mesh.close()
    out = {"workload": f"level-7 shell mesh, {len(blocks)} x 16^3 ({ncell} cells), Dirichlet outer faces, "
                       f"{steps} Jacobi steps (sweep + ghost exchange) in one bcm_jacobi_poisson_iterate call",
           "GCUPS": ncell * steps / el / 1e9, "ms_per_step": el / steps * 1e3,
           "halo_step_ms": havg, "first_and_last_step_ms": ems / max(en, 1), "closing_ghost_fill_ms": gms / max(gn, 1),
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles("jacobi_halo_n16"),
                        "note": "k_jacobi_halo<double,16,false>: 24 B per cell (compact iterate read, compact "
                                "right-hand side read, compact iterate written)"},
           "two_calls_per_step": {"ms_per_step": el2 / k * 1e3, "GCUPS": ncell * k / el2 / 1e9,
                                  "note": "bcm_jacobi7_poisson then bcm_ghost_fill per step on the ghosted field"}}
    print(f"[bench] jacobi level 7: {out['GCUPS']:.1f} GCUPS, halo step {havg:.3f} ms "
          f"({achieved:.0f} GB/s), two calls per step {el2 / k * 1e3:.3f} ms", file=sys.stderr, flush=True)
    return out


def true_relres(mesh, x, rhs_field, bvalues):
    """FP64 ||b - A x|| / ||b|| with the calc_r / assembly kernels (bit-exact with the oracle's)."""
    b = mesh.field(np.float64)
    mesh.poisson_assemble(None, b, rhs_field, bvalues)
    r = mesh.field(np.float64)
    mesh.calc_r_poisson(r, x, b)
    rr, bb = mesh.dot(r, r), mesh.dot(b, b)
    return float(np.sqrt(rr / bb))


def bench_bicgstab(args):
    """Mixed FP32/FP64 BiCGSTAB on the level-7 shell mesh (26,888 x 16^3) to 1e-10."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    tol, level, n = 1e-10, args.level, 16
    t = W.config3_tree(level)
    blocks = t.blocks
    ncell = len(blocks) * n ** 3
    print(f"[bench] bicgstab: level {level}, {len(blocks)} blocks x {n}^3 = {ncell} cells", file=sys.stderr, flush=True)
    f = W.rhs_field(blocks, n, 1, seed=3)
    rhs_host = -f
    del f
    mesh = Mesh(blocks, n, 1)
    mesh.set_progress(20.0)
    rhs = mesh.field(np.float64, rhs_host)
    vals = [0.0]

    def timed(fn, x):
        mesh.fill(x, 0.0)
        mesh.sync()
        t0 = time.perf_counter()
        out = fn(x)
        mesh.sync()
        return out, time.perf_counter() - t0

    x = mesh.field(np.float64)
    ug = not args.no_graph
    mi = args.max_iter
    mixed = lambda xx: mesh.solve(xx, rhs, vals, precision="mixed", tol=tol, max_iter=mi, use_graph=ug)[0]
    fp64 = lambda xx: mesh.solve(xx, rhs, vals, precision="fp64", tol=tol, max_iter=mi, use_graph=ug)[0]
    if args.precision != "auto":
        # One engine precision only (e.g. for a kernel trace of that path).
        return bicgstab_one(args, mesh, rhs, vals, timed, mixed if args.precision == "mixed" else fp64,
                            level, blocks, ncell, n)
    for _ in range(args.warmup):
        timed(mixed, x)
    times, st = [], None
    for _ in range(args.steps):
        st, dt = timed(mixed, x)
        times.append(dt)
        print(f"[bench] mixed solve: {st} {dt:.3f} s", file=sys.stderr, flush=True)
    xm = x.download()
    relres_mixed = true_relres(mesh, x, rhs, vals)
    x64 = mesh.field(np.float64)
    timed(fp64, x64)
    st64, t64 = timed(fp64, x64)
    print(f"[bench] fp64 engine solve: {st64} {t64:.3f} s", file=sys.stderr, flush=True)
    xh = mesh.field(np.float64)
    (rc, ith, rrh), th = timed(lambda xx: mesh.solve_fp64(xx, rhs, vals, tol=tol, max_iter=20000), xh)
    print(f"[bench] host-driven fp64 solve: rc={rc} iters={ith} relres={rrh:.3e} {th:.3f} s", file=sys.stderr,
          flush=True)
    a, b = W.interior(xm, 1), W.interior(x64.download(), 1)
    err = float(np.abs(a - b).max() / np.abs(b).max())
    elapsed = sum(times)
    solves_per_s = len(times) / elapsed
    res = {
        "metric": "solves/s (BiCGSTAB to FP64 relative residual 1e-10, mixed FP32/FP64)",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / len(times) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"adaptive shell octree level {level} (|x-c| = 0.25, margin 0.25), "
                               f"{len(blocks)} x 16^3 blocks, Jacobi(2 sweeps, 0.8)-preconditioned BiCGSTAB, "
                               "FP32 Krylov vectors / FP64 SpMV inputs, FP64 refinement to 1e-10",
                   "blocks": len(blocks), "interior_cells": ncell, "tol": tol},
        "mixed": {"iters": st["iters"], "refreshes": st["refreshes"], "relres": relres_mixed,
                  "err_vs_fp64_solution": err},
        "fp64_engine": {"iters": st64["iters"], "seconds": t64, "solves_per_s": 1.0 / t64,
                        "relres": st64["relres"]},
        "fp64_host_driven": {"iters": ith, "seconds": th, "solves_per_s": 1.0 / th, "relres": rrh},
        "speedup_vs_fp64_host_driven": th / (elapsed / len(times)),
        "speedup_vs_fp64_engine": t64 / (elapsed / len(times)),
    }
    # Algorithmic HBM bytes of one fused FP64 iteration, averaged over the
    # whole solve (reductions, graph launches and checks included).
    bpc = engine_bytes_per_cell(8, 2, n)
    achieved = bpc * ncell * st64["iters"] / t64 / 1e9
    tpc = traffic_from_profiles("bicgstab_fp64_n16_bytes_per_cell_iteration") if n == 16 else None
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                       "frac": achieved / 8000.0, "traffic": tpc * ncell if tpc else None,
                       "kernel": f"fused FP64 BiCGSTAB iteration, all kernels ({bpc:.0f} B per cell per iteration; "
                                 "traffic = measured HBM bytes per iteration)"}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_bicgstab(blocks, n, rhs_host, ith, args.cpu_steps)
    return res


def bicgstab_one(args, mesh, rhs, vals, timed, solve, level, blocks, ncell, n):
    x = mesh.field(np.float64)
    for _ in range(args.warmup):
        timed(solve, x)
    times, st = [], None
    for _ in range(args.steps):
        st, dt = timed(solve, x)
        times.append(dt)
        print(f"[bench] {args.precision} solve: {st} {dt:.3f} s", file=sys.stderr, flush=True)
    elapsed = sum(times)
    es = 8 if args.precision == "fp64" else 4
    res = {
        "metric": f"solves/s (BiCGSTAB to FP64 relative residual 1e-10, {args.precision} engine)",
        "value": len(times) / elapsed, "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / len(times) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"adaptive shell octree level {level} (|x-c| = 0.25, margin 0.25), "
                               f"{len(blocks)} x 16^3 blocks, Jacobi(2 sweeps, 0.8)-preconditioned BiCGSTAB",
                   "blocks": len(blocks), "interior_cells": ncell, "tol": 1e-10},
        "iters": st["iters"], "refreshes": st["refreshes"], "relres": true_relres(mesh, x, rhs, vals),
        "hip_graph": not args.no_graph,
    }
    if args.precision == "fp64":
        bpc = engine_bytes_per_cell(es, 2, n)
        achieved = bpc * ncell * st["iters"] / (elapsed / len(times)) / 1e9
        tpc = traffic_from_profiles("bicgstab_fp64_n16_bytes_per_cell_iteration") if n == 16 else None
        res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                           "frac": achieved / 8000.0, "traffic": tpc * ncell if tpc else None,
                           "kernel": f"fused FP64 BiCGSTAB iteration, all kernels ({bpc:.0f} B per cell per "
                                     "iteration; traffic = measured HBM bytes per iteration)"}
    return res


def engine_bytes_per_cell(es, pre, n):
    """Algorithmic bytes per interior cell of one iteration of the fused
    BiCGSTAB engine with vectors of `es` bytes (FP64: 8) and `pre` >= 1 Jacobi
    sweeps: every array read once and written once; a stencil input adds its
    4-column export (4/n of a vector).  With two sweeps and n <= 16 both
    sweeps run inside the p and s updates, which read the X-face frame from
    the exports of r, q and p."""
    xf = 4.0 / n
    d = 8.0
    if pre == 2 and n <= 16:
        upd_p = 3 * es * (1 + xf) + es * (1 + xf) + d * (1 + xf)  # r, q, p + exports -> p + export, p_ + export
        upd_s = 2 * es * (1 + xf) + es + d * (1 + xf)  # r, q + exports -> s, s_ + export
        spmv_q = d * (1 + xf) + es + es * (1 + xf)  # p_ + export, r0 -> q + export
        spmv_t = d * (1 + xf) + es + es  # s_ + export, s -> t
        xr = 3 * d + 3 * es + d + es * (1 + xf)  # dx, p_, s_, s, t, r0 -> dx, r + export
        return upd_p + upd_s + spmv_q + spmv_t + xr
    upd_p = 3 * es + es + es * (1 + xf)  # r, q, p -> p, first sweep + export
    upd_s = 2 * es + es + es * (1 + xf)  # r, q -> s, first sweep + export
    sweep = es * (1 + xf) + es + es * (1 + xf)  # input + export, rhs -> output + export
    spmv = d * (1 + xf) + es + es  # p_ + export, r0 (or s) -> q (or t)
    xr = 3 * d + 3 * es + d + es  # dx, p_, s_, s, t, r0 -> dx, r
    return upd_p + upd_s + 2 * (pre - 1) * sweep + 2 * spmv + xr


def cpu_bicgstab(blocks, n, rhs, gpu_iters, iters):
    """The oracle's ILS3D BiCGSTAB (FP64, OpenMP) for a few iterations on the
    same mesh; time-to-solution extrapolated to the GPU FP64 iteration count."""
    import oracle

    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, [0.0])
    x = np.zeros_like(rhs)
    work = [np.zeros_like(rhs) for _ in range(9)]
    t0 = time.perf_counter()
    oracle.bicgstab(blocks, n, 1, x, A, b, work, 0.8, 2, iters, 1e-30)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    per_it = dt / iters
    return {"value": 1.0 / (per_it * gpu_iters), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"same mesh, {iters} FP64 BiCGSTAB iterations ({dt:.1f} s, {per_it:.2f} s/iteration) on "
                      f"{cores} threads of {cpu_model()}, extrapolated to the GPU's {gpu_iters} iterations"}


def _batch_tables(M):
    from bcmtools_amd import workloads as W
    return [W.random_tree(1000 + i).blocks for i in range(M)]


def _build_batch(tables, idx, n=16):
    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    mesh = Mesh.batch([tables[i] for i in idx], n)
    mesh.set_progress(20.0)
    x, rhs = mesh.field(np.float64), mesh.field(np.float64)
    for s, i in enumerate(idx):
        first, cnt = mesh.segment(s)
        rhs.upload_blocks(first, -W.rhs_field_fast(tables[i], n, 1, seed=5000 + i))
    return mesh, x, rhs


def _timed_solve(mesh, x, rhs, precision, tol):
    mesh.fill(x, 0.0)
    mesh.sync()
    t0 = time.perf_counter()
    st = mesh.solve(x, rhs, [0.0], precision=precision, tol=tol, max_iter=5000)
    mesh.sync()
    return st, time.perf_counter() - t0


def _batch_summary(st):
    ok = sum(1 for e in st if e["status"] == "converged")
    return {"converged": ok, "segments": len(st), "max_relres": max(e["relres"] for e in st),
            "max_iters": max(e["iters"] for e in st), "sum_iters": sum(e["iters"] for e in st)}


def batched_subset(tables, every=8, tol=1e-10, n=16):
    """Every `every`-th mesh of the 256-mesh batch generator: solved as one
    batch (mixed and FP64), and one mesh at a time by the fused FP64 engine
    and by the plain host-driven FP64 path (ILS3D's BiCGSTAB step by step)."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    idx = list(range(0, len(tables), every))
    mesh, x, rhs = _build_batch(tables, idx, n)
    mesh.solve(x, rhs, [0.0], precision="mixed", tol=tol, max_iter=2)
    st_m, t_m = _timed_solve(mesh, x, rhs, "mixed", tol)
    mesh.solve(x, rhs, [0.0], precision="fp64", tol=tol, max_iter=2)
    st_d, t_d = _timed_solve(mesh, x, rhs, "fp64", tol)
    print(f"[bench] subset of {len(idx)}: batched mixed {t_m:.2f} s, batched fp64 {t_d:.2f} s", file=sys.stderr,
          flush=True)
    mesh.close()
    del x, rhs
    t_e = t_h = 0.0
    host_ok = 0
    for k, i in enumerate(idx):
        m1 = Mesh(tables[i], n)
        r1 = m1.field(np.float64, -W.rhs_field_fast(tables[i], n, 1, seed=5000 + i))
        xe, xh = m1.field(np.float64), m1.field(np.float64)
        m1.sync()
        t0 = time.perf_counter()
        m1.solve(xe, r1, [0.0], precision="fp64", tol=tol, max_iter=5000)
        m1.sync()
        t1 = time.perf_counter()
        rc, it, rr = m1.solve_fp64(xh, r1, [0.0], tol=tol, max_iter=5000)
        m1.sync()
        t2 = time.perf_counter()
        host_ok += rr <= tol
        t_e += t1 - t0
        t_h += t2 - t1
        m1.close()
        if k % 8 == 7:
            print(f"[bench] per-mesh solves: {k + 1}/{len(idx)} (fused fp64 {t_e:.2f} s, host-driven fp64 {t_h:.2f} s)",
                  file=sys.stderr, flush=True)
    ns = len(idx)
    return {"meshes": ns, "every": every, "blocks": int(sum(len(tables[i]) for i in idx)),
            "batched_mixed_solves_per_s": ns / t_m, "batched_fp64_solves_per_s": ns / t_d,
            "per_mesh_fused_fp64_solves_per_s": ns / t_e, "per_mesh_host_fp64_solves_per_s": ns / t_h,
            "speedup_batched_mixed_vs_per_mesh_host_fp64": t_h / t_m,
            "speedup_batched_mixed_vs_batched_fp64": t_d / t_m,
            "batched_mixed": _batch_summary(st_m), "batched_fp64": _batch_summary(st_d),
            "host_fp64_converged": int(host_ok)}


def bench_batched(args):
    """256 independent random octree meshes (16^3 blocks) solved together to 1e-10."""
    tol, n, M = 1e-10, 16, args.meshes
    tables = _batch_tables(M)
    counts = [len(tb) for tb in tables]
    levels = [int(tb["level"].max()) + 1 for tb in tables]
    total = sum(counts)
    print(f"[bench] batched: {M} meshes, {total} blocks ({min(counts)}..{max(counts)} per mesh)", file=sys.stderr,
          flush=True)
    # Comparison on a fixed subset (every 8th mesh); the faster batched
    # precision is then used for the full batch.
    sub = batched_subset(tables, 8, tol, n)
    t_m, t_d = 1.0 / sub["batched_mixed_solves_per_s"], 1.0 / sub["batched_fp64_solves_per_s"]
    precision = args.precision if args.precision != "auto" else ("mixed" if t_m < t_d else "fp64")

    # Full batch with that precision (the headline).
    mesh, x, rhs = _build_batch(tables, range(M), n)
    print(f"[bench] batched: fields uploaded, precision {precision}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        mesh.solve(x, rhs, [0.0], precision=precision, tol=tol, max_iter=2)
    times, st = [], None
    for _ in range(args.steps):
        st, dt = _timed_solve(mesh, x, rhs, precision, tol)
        times.append(dt)
        print(f"[bench] batched {precision} solve of {M} meshes: {dt:.2f} s {_batch_summary(st)}", file=sys.stderr,
              flush=True)
    full = _batch_summary(st)
    mesh.close()
    del x, rhs
    elapsed = sum(times)
    return {
        "metric": "batched solves/s (256 meshes, BiCGSTAB to FP64 relative residual 1e-10)",
        "value": M * len(times) / elapsed,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / len(times) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32+f64", "data": "synthetic",
        "precision": precision,
        "config": {"workload": f"{M} random octree meshes (max level U{{2..6}}, 1-3 sphere surfaces r~U(0.05,0.3), "
                               "2:1 balanced, 16^3 blocks) in one batch, Jacobi-preconditioned BiCGSTAB to 1e-10, "
                               "hipGraph-replayed iteration blocks, precision chosen on the every-8th subset",
                   "meshes": M, "blocks": total, "blocks_per_mesh": [min(counts), int(np.median(counts)), max(counts)],
                   "levels": [min(levels), max(levels)], "interior_cells": total * n ** 3},
        "batch": full,
        "subset": sub,
    }


def serve_stream(args, lam, trees, tols, gaps):
    """One pass of the request stream at arrival rate lam; returns the
    per-request outcomes and the stream's statistics."""
    from bcmtools_amd import Server

    R, n = len(trees), 16
    arrivals = np.cumsum(gaps / lam)
    sv = Server(200e9, max_active=args.slots, n=n, max_iter=20000, check_every=8,
                batch_blocks=getattr(args, "batch_blocks", 0), window_blocks=getattr(args, "window_blocks", 0),
                max_wait=getattr(args, "max_wait", 0.0))
    done, k = {}, 0
    t0 = sv.clock()
    last_print = t0
    while len(done) < R:
        now = sv.clock() - t0
        while k < R and arrivals[k] <= now:
            if args.policy == "spec":
                prec = "fp32" if tols[k] > 1e-8 else "mixed"
            else:
                prec = args.policy
            sv.submit(trees[k].blocks, float(tols[k]), prec, rhs=None, seed=9000 + k, tag=k)
            k += 1
        wait = (arrivals[k] - now) if k < R else 0.05
        sv.progress(max(0.0, min(0.01, wait)))
        for r in sv.poll():
            done[r["tag"]] = r
        if sv.clock() - last_print > 15.0:
            last_print = sv.clock()
            st = sv.stats()
            print(f"[bench] serving at {lam}/s: {len(done)}/{R} done, {st['active']} active, {st['queued']} queued, "
                  f"{st['bytes_in_use'] / 1e9:.1f} GB held", file=sys.stderr, flush=True)
    sv.close()
    t_end = max(r["t_done"] for r in done.values()) - t0
    lat = np.array([done[i]["t_done"] - t0 - arrivals[i] for i in range(R)])
    wait_q = np.array([done[i]["t_admit"] - t0 - arrivals[i] for i in range(R)])
    ok = [done[i]["status"] == 1 and done[i]["relres"] <= tols[i] for i in range(R)]
    failed = [{"tag": i, "status": done[i]["status"], "relres": done[i]["relres"], "iters": done[i]["iters"],
               "nblocks": done[i]["nblocks"], "tol": float(tols[i])} for i in range(R) if not ok[i]]
    by_tol = {}
    for tv in (1e-6, 1e-10):
        sel = [i for i in range(R) if tols[i] == tv]
        if sel:
            by_tol[f"{tv:g}"] = {"requests": len(sel), "p50_s": float(np.percentile(lat[sel], 50)),
                                 "p99_s": float(np.percentile(lat[sel], 99)),
                                 "mean_iters": float(np.mean([done[i]["iters"] for i in sel]))}
    span = t_end - arrivals[0]
    print(f"[bench] serving at {lam}/s: {R / span:.3f} solves/s, p50 {np.percentile(lat, 50):.2f} s, "
          f"p99 {np.percentile(lat, 99):.2f} s", file=sys.stderr, flush=True)
    return {"lambda_per_s": lam, "solves_per_s": R / span, "makespan_s": float(t_end),
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "mean": float(lat.mean()), "max": float(lat.max())},
            "queue_wait_s": {"p50": float(np.percentile(wait_q, 50)), "p99": float(np.percentile(wait_q, 99))},
            "by_tolerance": by_tol, "all_met_tolerance": bool(all(ok)), "not_met": failed}


def bench_serving(args):
    """Poisson-arrival stream of solves (block counts log-uniform in [64, 50k],
    tol 1e-6 or 1e-10), HBM budget 200 GB, at one or several arrival rates
    (the same requests, arrival gaps rescaled)."""
    from bcmtools_amd import workloads as W

    R = args.requests
    lams = [float(v) for v in args.lams.split(",")] if args.lams else [args.lam]
    rng = np.random.default_rng(args.seed)
    targets = np.exp(rng.uniform(np.log(64), np.log(args.max_blocks), size=R)).astype(int)
    tols = rng.choice([1e-6, 1e-10], size=R)
    gaps = rng.exponential(1.0, size=R)
    tb = time.perf_counter()
    trees = [W.serving_tree(7000 + i, int(targets[i])) for i in range(R)]
    t_tree = time.perf_counter() - tb
    counts = [len(t.blocks) for t in trees]
    print(f"[bench] serving: {R} requests, {sum(counts)} blocks ({min(counts)}..{max(counts)}), meshes built in "
          f"{t_tree:.2f} s on the host", file=sys.stderr, flush=True)
    runs = [serve_stream(args, lam, trees, tols, gaps) for lam in lams]
    head = runs[0]
    res = {
        "metric": "served Poisson solves/s with p50/p99 latency (continuous batching, HBM budget 200 GB)",
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1, "steps": R, "warmup": 0,
        "ms_per_step": 1e3 / head["solves_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.policy == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"{R} requests, Poisson arrivals at {lams[0]:g}/s, meshes of 1-3 random sphere "
                               f"surfaces with block counts log-uniform in [64, {args.max_blocks}] (16^3 blocks), "
                               "tol 1e-6 or 1e-10, "
                               + {"spec": "FP32 path for 1e-6 and FP32 + FP64 refinement for 1e-10",
                                  "mixed": "mixed FP32/FP64 BiCGSTAB (FP64 reliable updates) for every request",
                                  "fp64": "fused FP64 BiCGSTAB for every request"}[args.policy]
                               + f", up to {args.slots} groups on the device, HBM budget 200 GB"
                               + (f", requests of at most {args.batch_blocks} blocks merged into batched groups of "
                                  f"at most {args.batch_blocks} blocks" if getattr(args, "batch_blocks", 0) else
                                  ", one request per group")
                               + (f", smallest groups first within {args.window_blocks} blocks in flight "
                                  f"(a group held back at most {args.max_wait:g} s)"
                                  if getattr(args, "window_blocks", 0) else ", every group as soon as it is ready"),
                   "batch_blocks": int(getattr(args, "batch_blocks", 0)),
                   "window_blocks": int(getattr(args, "window_blocks", 0)),
                   "max_wait_s": float(getattr(args, "max_wait", 0.0)),
                   "precision_policy": args.policy,
                   "requests": R, "lambda_per_s": lams[0], "blocks_total": int(sum(counts)),
                   "blocks_per_request": [int(min(counts)), int(np.median(counts)), int(max(counts))],
                   "host_mesh_build_s_per_request": t_tree / R},
    }
    res.update({k: v for k, v in head.items() if k not in ("lambda_per_s", "solves_per_s")})
    if len(runs) > 1:
        res["sweep"] = [{k: r[k] for k in ("lambda_per_s", "solves_per_s", "makespan_s", "latency_s",
                                          "all_met_tolerance")} for r in runs]
    if args.per_request_every > 0:
        res["one_at_a_time"] = serve_baseline(trees, tols, args.per_request_every, head["solves_per_s"],
                                              np.cumsum(gaps / lams[0]))
    return res


def serve_baseline(trees, tols, every, served_rate, arrivals):
    """The requests of the stream solved alone, one after the other: every
    request by the fused FP64 engine, every `every`-th also by the plain
    host-driven FP64 path (bcm_solve_fp64, ILS3D's BiCGSTAB step by step).
    The plain path's time for the whole stream is estimated as the fused
    time times the plain/fused time ratio on the sampled requests; the
    latencies a first-come-first-served queue of one-at-a-time fused solves
    would give follow from the fused times and the arrival times.
    Right-hand sides of the served form (sin x sin y sin z + 0.01 N(0,1),
    negated) are drawn on the host."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    n = 16
    t_fused = t_fused_s = t_plain_s = 0.0
    ok, times = [], []
    for i, tr in enumerate(trees):
        blocks = tr.blocks
        mesh = Mesh(blocks, n, 1)
        rhs = mesh.field(np.float64, -W.rhs_field_fast(blocks, n, 1, seed=9000 + i))
        x = mesh.field(np.float64)
        mesh.sync()
        t0 = time.perf_counter()
        st = mesh.solve(x, rhs, [0.0], precision="fp64", tol=float(tols[i]), max_iter=20000)[0]
        mesh.sync()
        dt = time.perf_counter() - t0
        t_fused += dt
        times.append(dt)
        line = f"[bench] alone {i}: {len(blocks)} blocks, fused {st['iters']} its {dt:.2f} s"
        if i % every == 0:
            t_fused_s += dt
            x2 = mesh.field(np.float64)
            mesh.sync()
            t0 = time.perf_counter()
            _, it, rr = mesh.solve_fp64(x2, rhs, [0.0], tol=float(tols[i]), max_iter=5000)
            mesh.sync()
            dp = time.perf_counter() - t0
            t_plain_s += dp
            ok.append(rr <= tols[i])
            line += f", plain {it} its {dp:.2f} s"
            del x2
        print(line, file=sys.stderr, flush=True)
        del mesh, rhs, x
    R = len(trees)
    fused_rate = R / t_fused
    plain_est = t_fused * t_plain_s / t_fused_s
    lat, free = [], 0.0
    for i in range(R):
        free = max(free, float(arrivals[i])) + times[i]
        lat.append(free - float(arrivals[i]))
    return {"requests": R, "fused_fp64_one_at_a_time_s": t_fused, "fused_fp64_one_at_a_time_solves_per_s": fused_rate,
            "fifo_one_at_a_time_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "plain_sampled_every": every, "plain_sampled_all_met_tolerance": bool(all(ok)),
            "plain_over_fused_time_on_sample": t_plain_s / t_fused_s,
            "plain_fp64_host_driven_est_s": plain_est, "plain_fp64_host_driven_est_solves_per_s": R / plain_est,
            "served_vs_fused_one_at_a_time": served_rate / fused_rate,
            "served_vs_plain_est": served_rate / (R / plain_est)}


def bench_copy(args):
    """Plain device copy of one config-2-sized FP64 field (calibrates HBM counters)."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    t = W.config2_tree()
    mesh = Mesh(t.blocks, 32, 1)
    a, b = mesh.field(np.float64), mesh.field(np.float64)
    nbytes = len(t.blocks) * 34 ** 3 * 8
    for _ in range(args.warmup):
        mesh.copy(b, a)
    mesh.sync()
    mesh.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mesh.copy(b, a)
    mesh.sync()
    el = time.perf_counter() - t0
    ms, nl = mesh.kernel_time(L.K_BLAS1)
    gbs = 2.0 * nbytes / (ms / max(nl, 1) * 1e-3) / 1e9
    return {"metric": "device copy bandwidth (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "copy of one 3816 x 34^3 FP64 field", "bytes_per_copy": nbytes}}


def ghost_fill_bytes(counts, n):
    """Algorithmic bytes of one complete ghost fill (one ghost layer, FP64):
    every ghost cell written once and every source cell it needs read once.
    Same-level face: n^2 ghosts, one source each.  F2C subface: (n/2)^2
    coarse ghosts, eight fine sources each.  C2F face: n^2 fine ghosts from
    the (n/2+1)^2 x 2 coarse cells the trilinear stencils touch."""
    copy, f2c, c2f = counts["flat"], counts["f2c"], counts["c2f"]
    return 8 * (copy * 2 * n * n + f2c * 9 * (n // 2) ** 2 + c2f * (n * n + 2 * (n // 2 + 1) ** 2))


def host_exchange(mesh, f, args):
    """The facade's host-resident updateVC (the driver's arena is
    authoritative): upload the whole field, fill the ghosts, download it,
    from ordinary (pageable) host memory and from page-locked memory
    (bcm_host_alloc, what the facade allocates)."""
    import ctypes
    import time

    from bcmtools_amd import _lib as L

    lib = L.load()
    nbytes = int(np.prod(f.shape)) * f.dtype.itemsize
    out = {}
    pinned = ctypes.c_void_p()
    L.check(lib.bcm_host_alloc(nbytes, ctypes.byref(pinned)), "bcm_host_alloc")
    try:
        pageable = np.zeros(nbytes // 8)
        for label, ptr in (("pageable", pageable.ctypes.data), ("pinned", pinned.value)):
            times = []
            for _ in range(max(1, args.warmup) + max(1, min(args.steps, 5))):
                t0 = time.perf_counter()
                L.check(lib.bcm_field_upload(f.handle, ptr, nbytes), "bcm_field_upload")
                L.check(lib.bcm_ghost_fill(f.handle, L.GF_ALL), "bcm_ghost_fill")
                L.check(lib.bcm_field_download(f.handle, ptr, nbytes), "bcm_field_download")
                times.append(time.perf_counter() - t0)
            out[label] = 1e3 * min(times[max(1, args.warmup):])
    finally:
        lib.bcm_host_free(pinned)
    out["field_bytes"] = nbytes
    return out


def comm_exchange(mesh, blocks, f, args):
    """Every link of the mesh through face buffers, as a multi-GPU exchange
    would move them (bcm_comm_pack into one device buffer, bcm_comm_unpack
    from it): wall time of each call, its face list upload included."""
    import ctypes

    from bcmtools_amd import _lib as L

    lib = L.load()
    send, recv = [], []
    for b in range(len(blocks)):
        for fc in range(6):
            ld = int(blocks["nb_level_diff"][b][fc])
            ids = [int(v) for v in blocks["nb_id"][b][fc]]
            if ld == 0 and ids[0] >= 0:
                send.append((ids[0], fc ^ 1, 0, L.COMM_SAME))
                recv.append((b, fc, 0, L.COMM_SAME))
            elif ld == -1 and ids[0] >= 0:
                send.append((ids[0], fc ^ 1, int(blocks["nb_subface"][b][fc]), L.COMM_C2F))
                recv.append((b, fc, 0, L.COMM_C2F))
            elif ld == 1:
                for q in range(4):
                    if ids[q] >= 0:
                        send.append((ids[q], fc ^ 1, 0, L.COMM_F2C))
                        recv.append((b, fc, q, L.COMM_F2C))
    sizes = [int(lib.bcm_comm_face_size(f.handle, k)) for _, _, _, k in send]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    nbytes = int(np.sum(sizes)) * 8
    stage = mesh.field(np.float64)  # device memory for the buffer
    buf = lib.bcm_field_device_ptr(stage.handle)
    assert lib.bcm_field_bytes(stage.handle) >= nbytes
    fs = (L.CommFace * len(send))(*[L.CommFace(b, fc, q, k, int(o)) for (b, fc, q, k), o in zip(send, offs)])
    fr = (L.CommFace * len(recv))(*[L.CommFace(b, fc, q, k, int(o)) for (b, fc, q, k), o in zip(recv, offs)])
    tp, tu = [], []
    for _ in range(max(1, args.warmup) + max(1, min(args.steps, 10))):
        t0 = time.perf_counter()
        L.check(lib.bcm_comm_pack(f.handle, fs, len(send), buf, nbytes), "bcm_comm_pack")
        t1 = time.perf_counter()
        L.check(lib.bcm_comm_unpack(f.handle, fr, len(recv), buf, nbytes), "bcm_comm_unpack")
        tp.append(t1 - t0)
        tu.append(time.perf_counter() - t1)
    skip = max(1, args.warmup)
    return {"faces": len(send), "buffer_bytes": nbytes, "pack_ms": 1e3 * min(tp[skip:]),
            "unpack_ms": 1e3 * min(tu[skip:])}


def rank_buffers(mesh, blocks, f, args, ranks=8):
    """The mesh split over `ranks` ranks by the reference's Partition: the
    per-peer buffers of every rank (bcm_comm_plan: host time of all send
    and receive plans), then, rank by rank, the pack of its sends and the
    unpack of its receives on this GPU (wall time per call, face-list
    upload included; the slowest rank is reported)."""
    from bcmtools_amd import CommPlan, partition
    from bcmtools_amd import _lib as L

    lib = L.load()
    rank_of = partition(len(blocks), ranks)
    t0 = time.perf_counter()
    sends = [CommPlan(blocks, mesh.n, mesh.vc, np.float64, rank_of, r, L.COMM_SEND) for r in range(ranks)]
    recvs = [CommPlan(blocks, mesh.n, mesh.vc, np.float64, rank_of, r, L.COMM_RECV) for r in range(ranks)]
    plan_ms = 1e3 * (time.perf_counter() - t0)
    stage = mesh.field(np.float64)  # device memory for one rank's buffer
    buf = lib.bcm_field_device_ptr(stage.handle)
    assert lib.bcm_field_bytes(stage.handle) >= max(p.total_bytes for p in sends + recvs)
    from bcmtools_amd import CommList
    lists = [(CommList(mesh, sends[r].faces, sends[r].nfaces), CommList(mesh, recvs[r].faces, recvs[r].nfaces))
             for r in range(ranks)]
    worst_p, worst_u, worst_pl, worst_ul = 0.0, 0.0, 0.0, 0.0
    for r in range(ranks):
        tp, tu = [], []
        for _ in range(max(1, args.warmup) + max(1, min(args.steps, 10))):
            t1 = time.perf_counter()
            L.check(lib.bcm_comm_pack(f.handle, sends[r].faces, sends[r].nfaces, buf, sends[r].total_bytes),
                    "bcm_comm_pack")
            t2 = time.perf_counter()
            L.check(lib.bcm_comm_unpack(f.handle, recvs[r].faces, recvs[r].nfaces, buf, recvs[r].total_bytes),
                    "bcm_comm_unpack")
            tp.append(t2 - t1)
            tu.append(time.perf_counter() - t2)
        skip = max(1, args.warmup)
        worst_p = max(worst_p, 1e3 * min(tp[skip:]))
        worst_u = max(worst_u, 1e3 * min(tu[skip:]))
        # The same over the uploaded lists: enqueue, then wait for the stream.
        tp, tu = [], []
        for _ in range(max(1, args.warmup) + max(1, min(args.steps, 10))):
            t1 = time.perf_counter()
            lists[r][0].pack(f, buf, sends[r].total_bytes)
            mesh.sync()
            t2 = time.perf_counter()
            lists[r][1].unpack(f, buf, recvs[r].total_bytes)
            mesh.sync()
            tp.append(t2 - t1)
            tu.append(time.perf_counter() - t2)
        worst_pl = max(worst_pl, 1e3 * min(tp[skip:]))
        worst_ul = max(worst_ul, 1e3 * min(tu[skip:]))
    for ls, lr in lists:
        ls.close()
        lr.close()
    del stage
    return {"ranks": ranks, "plan_ms_all_ranks": plan_ms,
            "peers_per_rank": [len(p.peers) for p in sends],
            "send_bytes_per_rank": [p.total_bytes for p in sends],
            "faces_between_ranks": sum(p.nfaces for p in sends),
            "pack_ms_slowest_rank": worst_p, "unpack_ms_slowest_rank": worst_u,
            "uploaded_list_pack_ms_slowest_rank": worst_pl, "uploaded_list_unpack_ms_slowest_rank": worst_ul}


def bench_sorinner(args):
    """The sor drivers with nLoopInner > 1 (examples/sor2/Solver.cpp:
    106-109): `--inner` sweeps + in-block BC per block between exchanges,
    on the 3,816 x 32^3 shell mesh, FP64 (bcm_sor_iterate_inner; the fused
    sweeps over the stored ghosts, since the halo schedule cannot serve an
    inner loop).  A cell update is one cell of one sweep."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    omega, inner = 1.6, max(1, args.inner)
    t = W.config2_tree()
    blocks, n, vc = t.blocks, 32, 1
    ncell = len(blocks) * n ** 3
    s = -W.rhs_field(blocks, n, vc, seed=0)
    mesh = Mesh(blocks, n, vc)
    gf = mesh.field(np.float64)
    gs = mesh.field(np.float64, s)
    vals = [0.0]
    mesh.dirichlet_mirror(gf, vals)
    mesh.sor_iterate(gf, gs, omega, vals, max(1, args.warmup), inner)
    mesh.sync()
    mesh.reset_timing()
    mesh.set_timing(True)
    t0 = time.perf_counter()
    mesh.sor_iterate(gf, gs, omega, vals, args.steps, inner)
    mesh.sync()
    elapsed = time.perf_counter() - t0
    mesh.set_timing(False)
    sw_ms, sw_n = mesh.kernel_time(L.K_SOR)
    gh_ms, gh_n = mesh.kernel_time(L.K_GHOST)
    sweep = sw_ms / max(sw_n, 1)
    return {"metric": "stencil GCUPS (FP64 red-black SOR, nLoopInner sweeps per exchange)",
            "value": ncell * args.steps * inner / elapsed / 1e9, "unit": "GCUPS", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"shell octree level 5, 3816 x 32^3 blocks, sor2 RB-SOR omega=1.6, nLoopInner "
                                   f"{inner} (sweep + in-block Dirichlet) then ghost fill + Dirichlet per step",
                       "inner": inner},
            "sweep_ms": sweep, "sweep_algorithmic_GBps": 24.0 * ncell / (sweep * 1e-3) / 1e9,
            "ghost_fill_ms": gh_ms / max(gh_n, 1)}


def bench_ghost(args):
    """The reference's updateVC on its own (include/BlockManager.h:311-726,
    include/Scalar3DUpdater.h:190-347): the complete ghost fill of an FP64
    field with one ghost layer, on the 3,816 x 32^3 shell mesh and the
    26,888 x 16^3 shell mesh, the whole fill and each axis alone."""
    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    out = {}
    for name, tree, n in (("shell3816x32", W.config2_tree(), 32), ("shell26888x16", W.config3_tree(), 16)):
        mesh = Mesh(tree.blocks, n, 1)
        f = mesh.field(np.float64, W.rhs_field_fast(tree.blocks, n, 1, seed=3))
        counts = mesh.task_counts()
        nbytes = ghost_fill_bytes(counts, n)
        row = {"blocks": len(tree.blocks), "n": n, "tasks": counts,
               "algorithmic_bytes": nbytes}
        for label, mask in (("all", L.GF_ALL), ("x", L.GF_ALL | L.GF_X), ("y", L.GF_ALL | L.GF_Y),
                            ("z", L.GF_ALL | L.GF_Z), ("blockwise", L.GF_ALL | L.GF_BLOCKWISE)):
            for _ in range(args.warmup):
                mesh.update_vc(f, mask)
            mesh.sync()
            mesh.set_timing(True)
            mesh.reset_timing()
            for _ in range(args.steps):
                mesh.update_vc(f, mask)
            mesh.sync()
            ms, nl = mesh.kernel_time(L.K_GHOST)
            mesh.set_timing(False)
            row[f"ms_{label}"] = ms / max(nl, 1)
        row["algorithmic_GBps_all"] = nbytes / (row["ms_all"] * 1e-3) / 1e9
        if n == 32:
            row["host_resident_exchange_ms"] = host_exchange(mesh, f, args)
            row["pack_unpack_ms"] = comm_exchange(mesh, tree.blocks, f, args)
            row["rank_buffers"] = rank_buffers(mesh, tree.blocks, f, args)
        print(f"[bench] ghost fill {name}: {row}", file=sys.stderr, flush=True)
        out[name] = row
        mesh.close()
    first = out["shell3816x32"]
    res_cpu = None
    if not args.no_cpu:
        # The oracle's fill (OpenMP over the face tasks, the reference's
        # copyVCFromNeighbor loop) on the same field, a few repetitions.
        import oracle
        tree = W.config2_tree()
        host = W.rhs_field_fast(tree.blocks, 32, 1, seed=3)
        oracle.ghost_fill(tree.blocks, 32, 1, host)
        reps, t0 = 3, time.perf_counter()
        for _ in range(reps):
            oracle.ghost_fill(tree.blocks, 32, 1, host)
        dt = (time.perf_counter() - t0) / reps
        res_cpu = {"value": first["algorithmic_bytes"] / dt / 1e9, "unit": "GB/s", "cores": os.cpu_count() or 1,
                   "kind": "port", "ms": 1e3 * dt,
                   "sample": f"the oracle's complete fill of the same 3,816 x 32^3 FP64 field, {reps} repetitions"}
        print(f"[bench] ghost fill on the host (oracle): {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
    res = {"metric": "ghost fill (updateVC) throughput, algorithmic bytes", "value": first["algorithmic_GBps_all"],
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": first["ms_all"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "complete ghost fill of one FP64 field (one ghost layer) on the level-5 shell "
                                   "mesh (3,816 x 32^3) and the level-7 shell mesh (26,888 x 16^3)"},
            "meshes": out}
    if res_cpu:
        res["cpu_baseline"] = res_cpu
    return res


def bench_mesh(args):
    """Mesh construction: the octree, the block table with its neighbour
    information (BCMOctree + BlockFactory, Utils/src/BCMOctree.cpp:23-478,
    as bcm_tree_build) on one host thread and on all of them, and the
    device mesh (bcm_mesh_create: the face lists of BlockManager::
    prepareForVCUpdate, src/BlockManager.cpp:63-114, 471-648, and the device
    tables).  The reference's own construction (compiled in oracle/_ref)
    is timed on the same SimpleDivider trees as the CPU baseline."""
    import ctypes
    import subprocess
    import time

    from bcmtools_amd import Mesh, Tree
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    lib = L.load()

    def build_s(params):
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        L.check(lib.bcm_tree_build(ctypes.byref(params), ctypes.byref(h)), "bcm_tree_build")
        dt = time.perf_counter() - t0
        lib.bcm_tree_destroy(h)
        return dt

    def simple(level):
        return Tree(root_n=(2, 2, 2), divider=L.DIV_SIMPLE, min_level=1, max_level=level, ordering=L.ORDER_Z,
                    balance=0)

    ref = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "ref_harness")
    threads = os.cpu_count() or 1
    cases = (("serving_request", W.serving_tree(7000 + 7, 50000), None),
             ("shell_level7", W.shell_tree(7), None),
             ("simple_level6", simple(6), "simple 1 6 Z 2 2 2 dir 16 1"),
             ("simple_level7", simple(7), "simple 1 7 Z 2 2 2 dir 16 1"))
    out = {}
    for name, tree, ref_args in cases:
        row = {"blocks": len(tree.blocks)}
        for label, nt in (("one_thread", 1), ("all_threads", 0)):
            lib.bcm_tree_set_threads(nt)
            for _ in range(max(1, args.warmup)):
                build_s(tree.params)
            row[f"tree_ms_{label}"] = 1e3 * min(build_s(tree.params) for _ in range(max(1, args.steps)))
        lib.bcm_tree_set_threads(0)
        row["speedup"] = row["tree_ms_one_thread"] / row["tree_ms_all_threads"]
        times = []
        for _ in range(max(1, min(args.steps, 3))):
            t0 = time.perf_counter()
            m = Mesh(tree.blocks, 16, 1)
            m.sync()
            times.append(time.perf_counter() - t0)
            m.close()
        row["mesh_create_ms"] = 1e3 * min(times)
        if ref_args and os.path.exists(ref) and not args.no_cpu:
            # The reference's construction, one rank, best of three processes
            # (its BlockManager is a process-wide singleton).
            runs = []
            for _ in range(3):
                r = subprocess.run([ref, "meshtime"] + ref_args.split(), capture_output=True, text=True, timeout=300)
                if r.returncode != 0:
                    break
                f = r.stdout.split()
                runs.append({f[i]: float(f[i + 1]) for i in range(0, len(f) - 1, 2)})
            if len(runs) == 3:
                assert int(runs[0]["blocks"]) == len(tree.blocks)
                best = {k: min(x[k] for x in runs) for k in ("octree_ms", "blocks_ms", "facelists_ms")}
                row["reference"] = {**best, "tree_and_blocks_ms": best["octree_ms"] + best["blocks_ms"]}
                row["speedup_tree_vs_reference"] = row["reference"]["tree_and_blocks_ms"] / row["tree_ms_all_threads"]
                row["speedup_mesh_vs_reference"] = (row["reference"]["tree_and_blocks_ms"] + best["facelists_ms"]) / (
                    row["tree_ms_all_threads"] + row["mesh_create_ms"])
        print(f"[bench] mesh {name}: {row}", file=sys.stderr, flush=True)
        out[name] = row
    first = out["serving_request"]
    res = {"metric": "host mesh build time (octree, block table with neighbours)", "value": first["tree_ms_all_threads"],
           "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": first["tree_ms_all_threads"], "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "bcm_tree_build of the served stream's largest request, the level-7 shell mesh and "
                                  "two SimpleDivider meshes, on one host thread and on all of them; bcm_mesh_create",
                      "host_threads": threads},
           "meshes": out}
    refrow = out["simple_level6"].get("reference")
    if refrow:
        res["cpu_baseline"] = {"value": refrow["tree_and_blocks_ms"], "unit": "ms", "cores": 1, "kind": "reference",
                               "sample": "the reference's BCMOctree + BlockFactory::makeBlock loop on the SimpleDivider "
                                         "tree of levels 1-6 (219,080 blocks), best of 3 processes; its "
                                         "prepareForVCUpdate face lists took a further "
                                         f"{refrow['facelists_ms']:.0f} ms"}
    return res


def bench_sortol(args):
    """Config 2 at its own size: the sor2 outer iteration (examples/sor2/
    Solver.cpp:77-133) on the level-5 shell mesh, 3,816 x 32^3, run until
    relres ||s - lap_h f|| / ||s|| <= 1e-8 at omega = 1.4 (1.6 diverges with
    the reference's frozen ghosts), checked every --chunk sweeps.  The run
    can span several processes: --state DIR keeps the field (a Dump2 restart
    file of LocalScalar3D's format) and the sweep count, time and residual
    history between them; each process stops after --budget seconds."""
    import ctypes

    from bcmtools_amd import Mesh
    from bcmtools_amd import workloads as W

    n, omega, tol = 32, args.omega, 1e-8
    blocks = W.shell_tree(args.tol_level).blocks
    mesh = Mesh(blocks, n, 1)
    sf = mesh.field(np.float64, -W.rhs_field_fast(blocks, n, 1, seed=0))
    f = mesh.field(np.float64)
    state = {"sweeps": 0, "time_s": 0.0, "history": []}
    sdir = args.state
    if sdir and os.path.exists(os.path.join(sdir, "state.json")):
        state = json.load(open(os.path.join(sdir, "state.json")))
        f.load2(state["sweeps"], "f", directory=sdir)
        print(f"[bench] sor to tolerance: resumed at {state['sweeps']} sweeps", file=sys.stderr, flush=True)
    else:
        mesh.dirichlet_mirror(f, [0.0])
    r = ctypes.c_double()
    rr = state["history"][-1][1] if state["history"] else float("inf")
    mesh.sync()
    t0 = time.perf_counter()
    while rr > tol and time.perf_counter() - t0 < args.budget:
        mesh.sor_iterate(f, sf, omega, [0.0], args.chunk)
        state["sweeps"] += args.chunk
        if mesh.lib.bcm_sor_residual_noalloc(f.handle, sf.handle, ctypes.byref(r)) != 0:
            raise RuntimeError("bcm_sor_residual_noalloc failed")
        rr = r.value
        state["history"].append([state["sweeps"], rr])
        print(f"[bench] sor to tolerance: {state['sweeps']} sweeps, relres {rr:.4e}, "
              f"{state['time_s'] + time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if not np.isfinite(rr):
            break
    mesh.sync()
    state["time_s"] += time.perf_counter() - t0
    if sdir:
        os.makedirs(sdir, exist_ok=True)
        f.dump2(state["sweeps"], "f", directory=sdir)
        with open(os.path.join(sdir, "state.json"), "w") as fp:
            json.dump(state, fp)
    ncell = len(blocks) * n ** 3
    mesh.close()
    return {"metric": "time to relres 1e-8, FP64 red-black SOR (sor2 outer iteration)", "value": state["time_s"],
            "unit": "s", "n_gpus": 1, "steps": state["sweeps"], "warmup": 0,
            "ms_per_step": 1e3 * state["time_s"] / max(state["sweeps"], 1), "higher_is_better": False,
            "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"adaptive shell octree level {args.tol_level}, {len(blocks)} x {n}^3 blocks, "
                                   f"vc=1, sor2 RB-SOR omega={omega} + Dirichlet + ghost fill per sweep, relres "
                                   f"checked every {args.chunk} sweeps", "blocks": len(blocks), "omega": omega,
                       "tol": tol},
            "sweeps": state["sweeps"], "relres": rr, "converged": bool(rr <= tol),
            "GCUPS": ncell * state["sweeps"] / max(state["time_s"], 1e-9) / 1e9,
            "history": state["history"][:: max(1, len(state["history"]) // 50)]}


def bench_restart(args):
    """Restart files of SolverTest (LocalScalar3D::Dump2 / Load2) of an FP64
    field on the level-7 shell mesh, written from HBM and read back into it."""
    import shutil
    import tempfile

    from bcmtools_amd import Mesh
    from bcmtools_amd import _lib as L
    from bcmtools_amd import workloads as W

    level, n, vc = args.level, 16, 1
    t = W.config3_tree(level)
    blocks = t.blocks
    data = W.rhs_field_fast(blocks, n, vc, seed=11)
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(np.float64, data)
    g = mesh.field(np.float64)
    nbytes = data.nbytes
    print(f"[bench] restart: {len(blocks)} blocks, field {nbytes / 1e9:.2f} GB", file=sys.stderr, flush=True)
    d = tempfile.mkdtemp(prefix="bcm_restart_", dir="/tmp")
    try:
        for _ in range(args.warmup):
            f.dump2(0, "p", directory=d)
            g.load2(0, "p", directory=d)
        td, tl = [], []
        for s in range(args.steps):
            t0 = time.perf_counter()
            f.dump2(s, "p", directory=d)
            t1 = time.perf_counter()
            g.load2(s, "p", directory=d)
            mesh.sync()
            t2 = time.perf_counter()
            td.append(t1 - t0)
            tl.append(t2 - t1)
            os.remove(os.path.join(d, f"dump-p-00000-{s:010d}.bin"))
            print(f"[bench] restart step {s}: dump2 {t1 - t0:.3f} s, load2 {t2 - t1:.3f} s", file=sys.stderr,
                  flush=True)
        exact = bool(np.array_equal(g.download(), data))
        # PCIe reference rates: whole-field copies between pageable host
        # memory and the device (bcm_field_download / bcm_field_upload).
        host = np.empty_like(data)
        f.download()
        t0 = time.perf_counter()
        L.check(mesh.lib.bcm_field_download(f.handle, host.ctypes.data, host.nbytes), "bcm_field_download")
        d2h = nbytes / (time.perf_counter() - t0) / 1e9
        t0 = time.perf_counter()
        L.check(mesh.lib.bcm_field_upload(g.handle, host.ctypes.data, host.nbytes), "bcm_field_upload")
        h2d = nbytes / (time.perf_counter() - t0) / 1e9
        del host
        res = {
            "metric": "restart file throughput (LocalScalar3D Dump2 + Load2, GB/s of field data)",
            "value": 2 * nbytes * len(td) / (sum(td) + sum(tl)) / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": (sum(td) + sum(tl)) / len(td) * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"FP64 field on the level-{level} shell mesh, {len(blocks)} x 18^3 stored cells, "
                                   "one Dump2 file per step in /tmp (page cache, no fsync), read back by Load2",
                       "blocks": len(blocks), "bytes": nbytes},
            "dump2_GBps": nbytes * len(td) / sum(td) / 1e9,
            "load2_GBps": nbytes * len(tl) / sum(tl) / 1e9,
            "pcie_pageable_d2h_GBps": d2h, "pcie_pageable_h2d_GBps": h2d,
            "round_trip_bit_exact": exact,
        }
        if not args.no_cpu:
            from oracle import restart as R
            path = os.path.join(d, "cpu.bin")
            t0 = time.perf_counter()
            R.dump2(data, n, vc, path)
            tw = time.perf_counter() - t0
            t0 = time.perf_counter()
            R.load2(np.empty_like(data), n, vc, path)
            tr = time.perf_counter() - t0
            os.remove(path)
            res["cpu_baseline"] = {"value": 2 * nbytes / (tw + tr) / 1e9, "unit": "GB/s", "cores": 1,
                                   "kind": "port",
                                   "sample": "the same field written and read from host memory by the restatement "
                                             f"(write {nbytes / tw / 1e9:.2f} GB/s, read {nbytes / tr / 1e9:.2f} GB/s)"}
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="sor",
                    choices=["sor", "jacobi", "bicgstab", "batched", "serving", "copy", "restart", "ghost", "mesh",
                             "sortol", "sorinner"])
    ap.add_argument("--requests", type=int, default=32, help="serving: number of requests")
    ap.add_argument("--lam", type=float, default=10.0, help="serving: arrival rate (requests/s)")
    ap.add_argument("--lams", default="", help="serving: comma-separated arrival rates to sweep (same requests)")
    ap.add_argument("--slots", type=int, default=16, help="serving: requests on the device at once")
    ap.add_argument("--seed", type=int, default=1, help="serving: request stream seed")
    ap.add_argument("--policy", default="fp64", choices=["fp64", "mixed", "spec"],
                    help="serving: fp64 for every request, or FP32 (tol 1e-6) / mixed (1e-10)")
    ap.add_argument("--per-request-every", type=int, default=4,
                    help="serving: also solve every request alone on the fused FP64 path and every k-th on the "
                         "plain host-driven path (0: off)")
    ap.add_argument("--max-blocks", type=int, default=50000,
                    help="serving: block counts are log-uniform in [64, max-blocks]")
    ap.add_argument("--batch-blocks", type=int, default=0,
                    help="serving: merge requests of at most this many blocks into batched groups of at most "
                         "this many blocks (0: every request on its own mesh)")
    ap.add_argument("--window-blocks", type=int, default=8192,
                    help="serving: launch the smallest groups' iterations first, within this many blocks in flight "
                         "(0: every group as soon as it is ready)")
    ap.add_argument("--max-wait", type=float, default=5.0,
                    help="serving: with --window-blocks, the longest a group's iterations are held back (s; 0: no "
                         "limit)")
    ap.add_argument("--meshes", type=int, default=256, help="batched: number of meshes")
    ap.add_argument("--precision", default="auto", choices=["auto", "mixed", "fp64"],
                    help="batched: precision of the full batch (auto: the faster on the subset); "
                         "bicgstab: auto times mixed, FP64 and the host-driven path, else that engine only")
    ap.add_argument("--level", type=int, default=7, help="bicgstab: shell refinement level")
    ap.add_argument("--tol-level", type=int, default=5, help="sortol: shell refinement level")
    ap.add_argument("--omega", type=float, default=1.4, help="sortol: SOR weight")
    ap.add_argument("--chunk", type=int, default=2000, help="sortol: sweeps between residual checks")
    ap.add_argument("--budget", type=float, default=900.0, help="sortol: seconds this process runs")
    ap.add_argument("--state", default="", help="sortol: directory keeping the run across processes")
    ap.add_argument("--inner", type=int, default=3, help="sorinner: sweeps per exchange (nLoopInner)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-iter", type=int, default=20000,
                    help="bicgstab: iteration limit of the engine solves (a short limit for counter runs)")
    ap.add_argument("--quick", action="store_true",
                    help="sor: the headline line only (no tolerance, batched and serving sub-objects)")
    ap.add_argument("--no-graph", action="store_true",
                    help="bicgstab: launch the iteration kernels one by one instead of replaying a hipGraph")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    res = {"sor": bench_sor, "jacobi": bench_jacobi, "bicgstab": bench_bicgstab, "batched": bench_batched,
           "serving": bench_serving, "copy": bench_copy, "restart": bench_restart,
           "ghost": bench_ghost, "mesh": bench_mesh, "sortol": bench_sortol,
           "sorinner": bench_sorinner}[args.workload](args)
    if args.workload == "sor" and not args.quick:
        # The rest of the metric: SOR to a tolerance, batched solves/s and
        # served solves/s with p50/p99 latency, each on a bounded sample.
        res["sor_to_tolerance"] = sor_to_tolerance()
        sub = batched_subset(_batch_tables(256), every=8)
        res["batched"] = {"workload": "every 8th mesh of the 256-mesh batch generator (seeds 1000 + 8k), "
                                      "BiCGSTAB to 1e-10", **sub}
        # Served with the fused FP64 engine for every request (the fastest
        # policy on these meshes, and the precision of the one-at-a-time
        # comparison), then the same stream with the precision chosen by
        # tolerance (FP32 for 1e-6, mixed for 1e-10).
        sv_args = argparse.Namespace(**vars(args))
        sv_args.requests, sv_args.lam, sv_args.lams, sv_args.policy, sv_args.per_request_every = 12, 100.0, "", \
            "fp64", 4
        sv = bench_serving(sv_args)
        keys = ("value", "unit", "config", "latency_s", "queue_wait_s", "by_tolerance", "all_met_tolerance",
                "not_met", "makespan_s")
        res["serving"] = {k: sv[k] for k in keys + ("one_at_a_time",) if k in sv}
        base = sv.get("one_at_a_time", {})
        if base.get("plain_fp64_host_driven_est_solves_per_s"):
            res["serving"]["speedup_vs_plain_fp64_per_request"] = \
                sv["value"] / base["plain_fp64_host_driven_est_solves_per_s"]
        sv_args.policy, sv_args.per_request_every = "spec", 0
        sp = bench_serving(sv_args)
        res["serving"]["precision_by_tolerance"] = {k: sp[k] for k in ("value", "unit", "latency_s", "by_tolerance",
                                                                       "all_met_tolerance", "makespan_s")}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
