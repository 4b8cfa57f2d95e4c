"""Write the golden fixtures of tests/golden/ from the reference itself.

Runs oracle/_ref/ref_harness -- the reference BCMTools library sources
compiled by oracle/ref/Makefile with our fixture-writer main -- and stores
its block tables and fields as .npz.  Only runs where the reference tree
exists (this container); the fixtures are committed and the tests read
them everywhere.

usage: python tests/golden/make_golden.py [fixture names: all when none]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from bcmtools_amd._lib import BLOCK_DTYPE  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF = "/root/reference"

MESHES = {
    # name: mesh args (divider minLevel maxLevel ordering rnx rny rnz bc n vc [sphere])
    "mesh_flat2_z": "flat 0 2 Z 1 1 1 dir 16 1",
    "mesh_simple13_z": "simple 1 3 Z 1 1 1 dir 4 1",
    "mesh_simple03_hilbert_x": "simple 0 3 Hilbert 1 1 1 x 2 1",
    "mesh_flat1_hilbert_roots_y": "flat 0 1 Hilbert 2 2 2 y 4 1",
    "mesh_simple24_hilbert_z": "simple 2 4 Hilbert 1 2 1 z 4 1",
}

GHOSTS = {
    # name: (method, dtype, mesh args)
    "ghost_sphere_atonce_d": (0, "d", "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3"),
    "ghost_sphere_xyz_d": (1, "d", "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3"),
    "ghost_sphere_leveldiff_d": (2, "d", "sphere 1 3 Hilbert 2 1 1 dir 4 1 1.0 0.5 0.5 0.3"),
    "ghost_sphere_atonce_f": (0, "f", "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3"),
    "ghost_simple_vc2_d": (0, "d", "simple 1 2 Z 1 1 1 dir 8 2"),
    "ghost_periodic_x_d": (0, "d", "simple 1 3 Z 1 1 1 x 4 1"),
    # two-cell ghost frames across level differences (F2C sums and C2F
    # interpolation reach two cells deep)
    "ghost_sphere_vc2_d": (0, "d", "sphere 1 3 Z 1 1 1 dir 4 2 0.5 0.5 0.5 0.3"),
    "ghost_sphere_vc2_leveldiff_f": (2, "f", "sphere 1 3 Hilbert 1 1 1 dir 4 2 0.5 0.5 0.5 0.3"),
}

VTK = {
    # name: (step, difflevel, mesh args); the .vthb index and every 16th
    # block's .vti (plus the last) are kept.
    "vtk_sphere": (7, 2, "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3"),
}

COMMRANK = {
    # name: (ranks, method, dtype, mesh args, ghost fixture of the same mesh,
    #        method and dtype, or None to store the mesh and fields here).
    # The reference's exchange on that many MPI ranks (mpiexec); every
    # message a rank sends is recorded through the MPI profiling interface.
    "comm3_sphere_atonce_d": (3, 0, "d", "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3", "ghost_sphere_atonce_d"),
    "comm2_sphere_xyz_d": (2, 1, "d", "sphere 1 3 Z 2 1 1 dir 4 1 1.0 0.5 0.5 0.3", "ghost_sphere_xyz_d"),
    "comm4_sphere_vc2_leveldiff_f": (4, 2, "f", "sphere 1 3 Hilbert 1 1 1 dir 4 2 0.5 0.5 0.5 0.3",
                                     "ghost_sphere_vc2_leveldiff_f"),
    "comm5_periodic_x_d": (5, 0, "d", "simple 1 3 Z 1 1 1 x 4 1", "ghost_periodic_x_d"),
    # n = 2 in FP32: 4-byte F2C slots, padded to 8 in the buffers
    "comm3_simple_n2_f": (3, 0, "f", "simple 0 3 Z 1 1 1 x 2 1", None),
}
MPIEXEC = "/opt/conda/bin/mpiexec"

SOR1 = {
    # name: (config text, roots, ordering)
    "sor1_log4": (None, 1, "Z"),
    "sor1_adaptive_y": ("minLevel = 1\nmaxLevel = 3\ntreeType = simple\nordering = Z\ntype = y\n"
                        "nLoopOuter = 25\nnLoopInner = 1\nomega = 1.5\nsize = 8\nvc = 1\nb0 = 0.25\nb1 = 2.0\n"
                        "output = out.dat\n", 1, "Z"),
    # nLoopInner = 3: three sweeps + in-block BC per block between exchanges
    "sor1_inner3_z": ("minLevel = 1\nmaxLevel = 3\ntreeType = simple\nordering = Hilbert\ntype = z\n"
                      "nLoopOuter = 12\nnLoopInner = 3\nomega = 1.3\nsize = 4\nvc = 1\nb0 = -0.5\nb1 = 1.25\n"
                      "output = out.dat\n", 1, "Hilbert"),
}


def read_dump(path, nfields, dtype):
    with open(path, "rb") as fp:
        hdr = np.frombuffer(fp.read(16), dtype=np.int32)
        nb, n, vc, rec = (int(v) for v in hdr)
        assert rec == BLOCK_DTYPE.itemsize, "block record size mismatch"
        blocks = np.frombuffer(fp.read(rec * nb), dtype=BLOCK_DTYPE).copy()
        nx0 = n + 2 * vc
        fields = []
        for _ in range(nfields):
            cnt = nb * nx0 ** 3
            fields.append(np.frombuffer(fp.read(cnt * np.dtype(dtype).itemsize), dtype=dtype)
                          .reshape(nb, nx0, nx0, nx0).copy())
    return blocks, n, vc, fields


def run(args, cwd):
    out = subprocess.run([HARNESS] + args, cwd=cwd, check=True, capture_output=True, text=True)
    return out.stdout


def main(only=()):
    def wanted(name):
        return not only or name in only

    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are only regenerated where it is")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, margs in MESHES.items():
            if not wanted(name):
                continue
            out = os.path.join(tmp, name + ".bin")
            run(["mesh", out] + margs.split(), tmp)
            blocks, n, vc, _ = read_dump(out, 0, np.float64)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), blocks=blocks.view(np.uint8), n=n, vc=vc,
                                args=margs)
            print(name, len(blocks))
        for name, (method, dt, margs) in GHOSTS.items():
            if not wanted(name):
                continue
            out = os.path.join(tmp, name + ".bin")
            run(["ghost", out, str(method), dt] + margs.split(), tmp)
            dtype = np.float64 if dt == "d" else np.float32
            blocks, n, vc, (fin, fout) = read_dump(out, 2, dtype)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), blocks=blocks.view(np.uint8), n=n, vc=vc,
                                method=method, fin=fin, fout=fout, args=margs)
            print(name, len(blocks), fin.shape)
        for name, (step, difflevel, margs) in VTK.items():
            if not wanted(name):
                continue
            d = os.path.join(tmp, name)
            os.makedirs(d)
            out = os.path.join(d, name + ".bin")
            run(["vtk", out, str(step), str(difflevel)] + margs.split(), d)
            blocks, n, vc, (fin,) = read_dump(out, 1, np.float64)
            vtkdir = os.path.join(d, "VTK")
            with open(os.path.join(vtkdir, f"data-p-{step:010d}.vthb"), "rb") as fp:
                vthb = np.frombuffer(fp.read(), dtype=np.uint8)
            keep = sorted(set(range(0, len(blocks), 16)) | {len(blocks) - 1})
            vti = {}
            for b in keep:
                with open(os.path.join(vtkdir, f"{step:010d}", f"data-p-00000-{b:05d}-{step:010d}.vti"), "rb") as fp:
                    vti[f"vti_{b}"] = np.frombuffer(fp.read(), dtype=np.uint8)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), blocks=blocks.view(np.uint8), n=n, vc=vc,
                                fin_kept=fin[keep], step=step, difflevel=difflevel, vthb=vthb, kept=np.array(keep),
                                args=margs,
                                **vti)
            print(name, len(blocks), "vti kept", len(keep))
        for name, (nprocs, method, dt, margs, ghost) in COMMRANK.items():
            if not wanted(name):
                continue
            dtype = np.float64 if dt == "d" else np.float32
            single = os.path.join(tmp, name + ".bin")
            run(["ghost", single, str(method), dt] + margs.split(), tmp)
            blocks, n, vc, (fin, fout) = read_dump(single, 2, dtype)
            prefix = os.path.join(tmp, name + ".rank")
            subprocess.run([MPIEXEC, "-n", str(nprocs), HARNESS, "commrank", prefix, str(method), dt] + margs.split(),
                           cwd=tmp, check=True, capture_output=True, text=True)
            src, dst, tag, start, end, lens, blobs, parts = [], [], [], [], [], [], [], []
            nx0 = n + 2 * vc
            for r in range(nprocs):
                with open(f"{prefix}.{r}", "rb") as fp:
                    rank, npr, st, en, nmsg = (int(v) for v in np.frombuffer(fp.read(20), dtype=np.int32))
                    assert (rank, npr) == (r, nprocs)
                    start.append(st)
                    end.append(en)
                    for _ in range(nmsg):
                        d, t = (int(v) for v in np.frombuffer(fp.read(8), dtype=np.int32))
                        nb = int(np.frombuffer(fp.read(8), dtype=np.int64)[0])
                        src.append(r)
                        dst.append(d)
                        tag.append(t)
                        lens.append(nb)
                        blobs.append(fp.read(nb))
                    parts.append(np.frombuffer(fp.read(), dtype=dtype).reshape(en - st, nx0, nx0, nx0))
            # The exchange over ranks gives the one-rank result.
            assert np.array_equal(np.concatenate(parts), fout), name
            extra = {}
            if ghost is None:
                extra = dict(blocks=blocks.view(np.uint8), n=n, vc=vc, fin=fin, fout=fout, args=margs)
            else:
                g = np.load(os.path.join(HERE, ghost + ".npz"))
                assert np.array_equal(g["fin"], fin) and np.array_equal(g["fout"], fout), (name, ghost)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), ranks=nprocs, method=method, dtype=dt,
                                ghost=ghost or "", start=np.array(start), end=np.array(end), msg_src=np.array(src),
                                msg_dst=np.array(dst), msg_tag=np.array(tag), msg_len=np.array(lens, dtype=np.int64),
                                msg_bytes=np.frombuffer(b"".join(blobs), dtype=np.uint8), **extra)
            print(name, nprocs, "ranks", len(src), "messages", sum(lens), "bytes")
        for name, (conf, rn, ordering) in SOR1.items():
            if not wanted(name):
                continue
            cfile = os.path.join(tmp, name + ".conf")
            if conf is None:
                with open(os.path.join(REF, "examples", "sor1", "test.conf")) as fp:
                    conf = fp.read()
            with open(cfile, "w") as fp:
                fp.write(conf)
            out = os.path.join(tmp, name + ".bin")
            text = run(["sor1", out, cfile, str(rn), ordering], tmp)
            err = float([ln for ln in text.splitlines() if ln.startswith("errorMax")][-1].split("=")[1])
            blocks, n, vc, (fout,) = read_dump(out, 1, np.float64)
            keys = {}
            for ln in conf.splitlines():
                if "=" in ln and not ln.strip().startswith("#"):
                    k, v = ln.split("=", 1)
                    keys[k.strip()] = v.strip()
            np.savez_compressed(os.path.join(HERE, name + ".npz"), blocks=blocks.view(np.uint8), n=n, vc=vc,
                                fout=fout, error_max=err, omega=float(keys["omega"]),
                                iters=int(keys["nLoopOuter"]), inner=int(keys.get("nLoopInner", 1)),
                                type=keys["type"],
                                b0=float(keys.get("b0", 0.0)), b1=float(keys.get("b1", 1.0)))
            print(name, len(blocks), "errorMax", err)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
