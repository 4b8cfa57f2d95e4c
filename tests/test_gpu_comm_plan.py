"""Per-peer buffers of one rank on the GPU: bcm_comm_plan + bcm_comm_pack /
bcm_comm_unpack against the reference's MPI messages.

Every rank of the recorded multi-rank exchange is played on one GPU in
turn, call by call (AtOnce; SeparateXYZ; SeparateXYZ_SeparateLevelDiff):
- pack: the rank's buffer holds, peer by peer, the bytes the reference
  rank sent that peer (every slot, padding aside);
- unpack: the messages the reference sent to the rank, placed at the
  receive plan's offsets and unpacked, write exactly the ghost boxes of
  the rank's faces toward other ranks, with the values of the one-GPU
  exchange;
- between calls the one-GPU exchange of that call advances the field, as
  the ranks' own local copies and receives would.
"""
import sys

import numpy as np
import pytest

from bcmtools_amd import CommList, CommPlan, Mesh, partition
from bcmtools_amd import _lib as L

from hipmem import DeviceArray
from test_comm_plan import COMM, calls, load_comm, split_by_call
from test_gpu_comm import recv_box

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def slot_count(n, vc, kind):
    full = n * n * vc
    return full // 4 if kind == L.COMM_F2C else full


def run_call(mesh, fld, axis, kind):
    if axis < 0:
        mesh.update_vc(fld)
    else:
        mesh.update_vc_xyz(fld, axis, kind)


@pytest.mark.parametrize("name", COMM)
def test_rank_buffers_carry_the_references_messages(name):
    c = load_comm(name)
    g = c["mesh"]
    blocks, n, vc = g["blocks"], g["n"], g["vc"]
    dtype = np.float64 if str(c["dtype"]) == "d" else np.float32
    es = np.dtype(dtype).itemsize
    ranks = int(c["ranks"])
    rank_of = partition(len(blocks), ranks)
    per_call = split_by_call(c, rank_of, dtype)
    case(f"comm plan {name}: {ranks} ranks, method {int(c['method'])}, nb={len(blocks)} n={n} vc={vc}")
    mesh = Mesh(blocks, n, vc)
    fld = mesh.field(dtype, g["fin"])
    for ci, (axis, sl, rl, kind) in enumerate(calls(int(c["method"]))):
        before = fld.download()
        unpacked = mesh.field(dtype, before)
        covered = np.zeros(before.shape, bool)
        for r in range(ranks):
            send = CommPlan(blocks, n, vc, dtype, rank_of, r, L.COMM_SEND, axis, sl)
            if send.nfaces:
                buf = DeviceArray(np.zeros(send.total_bytes, np.uint8))
                assert mesh.lib.bcm_comm_pack(fld.handle, send.faces, send.nfaces, buf.ptr, send.total_bytes) == L.OK
                host = buf.get()
                # The same faces through a list uploaded once.
                buf.put(np.zeros(send.total_bytes, np.uint8))
                lst = CommList(mesh, send.faces, send.nfaces)
                lst.pack(fld, buf.ptr, send.total_bytes)
                mesh.sync()
                np.testing.assert_array_equal(buf.get(), host)
                lst.close()
                buf.free()
                for (peer, first, nf, off, nbytes), (dest, msg) in zip(send.peers, per_call[r][ci]):
                    assert peer == dest and nbytes == len(msg)
                    for f in send.faces[first:first + nf]:
                        a = f.offset * es - off
                        b = a + slot_count(n, vc, f.kind) * es
                        np.testing.assert_array_equal(host[off + a:off + b], msg[a:b])
            recv = CommPlan(blocks, n, vc, dtype, rank_of, r, L.COMM_RECV, axis, rl)
            if recv.nfaces:
                host = np.zeros(recv.total_bytes, np.uint8)
                sent = {s: msg for s in range(ranks) for d, msg in per_call[s][ci] if d == r}
                for peer, first, nf, off, nbytes in recv.peers:
                    host[off:off + nbytes] = sent[peer]
                buf = DeviceArray(host)
                assert mesh.lib.bcm_comm_unpack(unpacked.handle, recv.faces, recv.nfaces, buf.ptr,
                                                recv.total_bytes) == L.OK
                buf.free()
                for f in recv.faces[:recv.nfaces]:
                    box = recv_box(covered[f.block], n, vc, f.face, f.subface, f.kind)
                    assert not box.any()  # each ghost cell has one sender
                    recv_box_view(covered[f.block], n, vc, f.face, f.subface, f.kind)[...] = True
        run_call(mesh, fld, axis, kind)
        after = fld.download()
        got = unpacked.download()
        np.testing.assert_array_equal(got[covered], after[covered])
        np.testing.assert_array_equal(got[~covered], before[~covered])
    np.testing.assert_array_equal(fld.download(), g["fout"])


def recv_box_view(a, n, vc, face, sub, kind):
    """Writable view of the ghost box recv_box reads."""
    axis, plus = face // 2, face % 2
    t0, t1 = (axis + 1) % 3, (axis + 2) % 3
    lo, ext = [0, 0, 0], [n, n, n]
    lo[axis], ext[axis] = (n if plus else -vc), vc
    if kind == L.COMM_F2C:
        h = n // 2
        ext[t0] = ext[t1] = h
        lo[t0], lo[t1] = h * (sub & 1), h * (sub >> 1)
    return a[tuple(slice(lo[d] + vc, lo[d] + vc + ext[d]) for d in (2, 1, 0))]


def test_eight_ranks_on_the_config2_mesh():
    """The 3,816 x 32^3 shell mesh over 8 ranks: every rank packs its sends,
    the buffers are routed peer to peer, every rank unpacks; the ghost boxes
    between ranks equal the one-GPU exchange, bit for bit."""
    from bcmtools_amd import workloads as W
    blocks = W.config2_tree().blocks
    n, vc, ranks = 32, 1, 8
    rank_of = partition(len(blocks), ranks)
    sends = [CommPlan(blocks, n, vc, np.float64, rank_of, r, L.COMM_SEND) for r in range(ranks)]
    recvs = [CommPlan(blocks, n, vc, np.float64, rank_of, r, L.COMM_RECV) for r in range(ranks)]
    case(f"comm plan config 2: {len(blocks)} blocks, {ranks} ranks, "
         f"{sum(p.nfaces for p in sends)} faces between ranks")
    rng = np.random.default_rng(8)
    data = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2))
    mesh = Mesh(blocks, n, vc)
    src = mesh.field(np.float64, data)
    out = mesh.field(np.float64, data)
    bufs = []
    for r in range(ranks):
        b = DeviceArray(np.zeros(max(8, sends[r].total_bytes), np.uint8))
        assert mesh.lib.bcm_comm_pack(src.handle, sends[r].faces, sends[r].nfaces, b.ptr, sends[r].total_bytes) == L.OK
        bufs.append(b.get())
        b.free()
    covered = np.zeros(data.shape, bool)
    for r in range(ranks):
        host = np.zeros(recvs[r].total_bytes, np.uint8)
        for peer, first, nf, off, nbytes in recvs[r].peers:
            sp = [p for p in sends[peer].peers if p[0] == r][0]
            assert sp[4] == nbytes
            host[off:off + nbytes] = bufs[peer][sp[3]:sp[3] + nbytes]
        b = DeviceArray(host)
        assert mesh.lib.bcm_comm_unpack(out.handle, recvs[r].faces, recvs[r].nfaces, b.ptr,
                                        recvs[r].total_bytes) == L.OK
        b.free()
        for f in recvs[r].faces[:recvs[r].nfaces]:
            recv_box_view(covered[f.block], n, vc, f.face, f.subface, f.kind)[...] = True
    mesh.update_vc(src)
    want = src.download()
    got = out.download()
    np.testing.assert_array_equal(got[covered], want[covered])
    np.testing.assert_array_equal(got[~covered], data[~covered])


def test_uploaded_face_lists_equal_the_per_call_lists():
    """bcm_comm_pack_list / bcm_comm_unpack_list over lists uploaded once:
    the buffers and fields of bcm_comm_pack / bcm_comm_unpack, bit for bit,
    over three exchanges of changing data on the 8-rank config-2 plans; a
    list is refused for another mesh's field and for a short buffer."""
    from bcmtools_amd import workloads as W
    blocks = W.config2_tree().blocks
    n, vc, ranks = 32, 1, 8
    rank_of = partition(len(blocks), ranks)
    mesh = Mesh(blocks, n, vc)
    case("comm lists: config 2, 8 ranks, 3 exchanges")
    plans = [(CommPlan(blocks, n, vc, np.float64, rank_of, r, L.COMM_SEND),
              CommPlan(blocks, n, vc, np.float64, rank_of, r, L.COMM_RECV)) for r in (0, 3, 7)]
    lists = [(CommList(mesh, s.faces, s.nfaces), CommList(mesh, q.faces, q.nfaces)) for s, q in plans]
    rng = np.random.default_rng(11)
    for it in range(3):
        data = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2))
        a = mesh.field(np.float64, data)
        b = mesh.field(np.float64, data)
        for (send, recv), (ls, lr) in zip(plans, lists):
            nb = max(send.total_bytes, recv.total_bytes)
            b1 = DeviceArray(np.zeros(nb, np.uint8))
            b2 = DeviceArray(np.zeros(nb, np.uint8))
            assert mesh.lib.bcm_comm_pack(a.handle, send.faces, send.nfaces, b1.ptr, send.total_bytes) == L.OK
            ls.pack(b, b2.ptr, send.total_bytes)
            mesh.sync()
            np.testing.assert_array_equal(b1.get(), b2.get())
            # Unpack the same received bytes both ways.
            msg = rng.standard_normal(nb // 8).view(np.uint8)
            b1.put(msg)
            b2.put(msg)
            assert mesh.lib.bcm_comm_unpack(a.handle, recv.faces, recv.nfaces, b1.ptr, recv.total_bytes) == L.OK
            lr.unpack(b, b2.ptr, recv.total_bytes)
            mesh.sync()
            b1.free()
            b2.free()
        np.testing.assert_array_equal(a.download(), b.download())
    other = Mesh(blocks, n, vc)
    f = other.field(np.float64)
    ls = lists[0][0]
    buf = DeviceArray(np.zeros(plans[0][0].total_bytes, np.uint8))
    assert mesh.lib.bcm_comm_pack_list(f.handle, ls.handle, buf.ptr, plans[0][0].total_bytes) == L.ERR_ARG
    g = mesh.field(np.float64)
    assert mesh.lib.bcm_comm_pack_list(g.handle, ls.handle, buf.ptr, plans[0][0].total_bytes - 8) == L.ERR_ARG
    assert mesh.lib.bcm_comm_pack_list(g.handle, ls.handle, buf.ptr, plans[0][0].total_bytes) == L.OK
    mesh.sync()
    buf.free()
    for ls, lr in lists:
        ls.close()
        lr.close()
