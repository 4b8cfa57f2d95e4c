"""Per-peer buffers of one rank (bcm_comm_plan) against the reference's own
MPI messages.

The fixtures hold every message each rank of the reference handed to
MPI_Isend during one exchange on 2-5 MPI ranks (the compiled reference
library under mpiexec, recorded through the MPI profiling interface; the
ranks' combined result equals the one-rank exchange).  Here, with no GPU:
for every rank and every exchange call, the plan names the same peers in
the same order with the same message sizes (CommBuffer: one buffer per
peer, slots padded to 8 bytes, src/CommBuffer.cpp:55-94), for the AtOnce,
SeparateXYZ and SeparateXYZ_SeparateLevelDiff face lists, FP64 and FP32,
and the receive plans agree with what the senders sent.  The bytes
themselves are compared on the GPU (test_gpu_comm_plan.py).
"""
import os

import numpy as np
import pytest

from bcmtools_amd import CommPlan, partition
from bcmtools_amd import _lib as L

from conftest import GOLDEN, load_golden

COMM = ["comm3_sphere_atonce_d", "comm2_sphere_xyz_d", "comm4_sphere_vc2_leveldiff_f", "comm5_periodic_x_d",
        "comm3_simple_n2_f"]


def load_comm(name):
    """The fixture, its mesh (blocks, n, vc, fin, fout; from the ghost
    fixture it names, or its own) and the messages per sending rank."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    c = {k: z[k] for k in z.files}
    ghost = str(c["ghost"])
    if ghost:
        g = load_golden(ghost)
    else:
        g = dict(c)
        g["blocks"] = c["blocks"].view(L.BLOCK_DTYPE).copy()
        g["n"], g["vc"] = int(c["n"]), int(c["vc"])
    c["mesh"] = g
    ranks = int(c["ranks"])
    msgs = [[] for _ in range(ranks)]
    pos = 0
    blob = c["msg_bytes"]
    for s, d, ln in zip(c["msg_src"], c["msg_dst"], c["msg_len"]):
        msgs[int(s)].append((int(d), blob[pos:pos + int(ln)]))
        pos += int(ln)
    c["msgs"] = msgs
    return c


def calls(method):
    """(axis, send level difference, receive level difference, ghost-fill
    kinds) of each exchange call of an update (BlockManager::updateVC,
    updateVC_X/Y/Z, and the F2C / Flat / C2F order of
    examples/SolverTest/LocalScalar3D.h:214-224)."""
    any_ = L.COMM_ANY_LEVEL
    if method == 0:
        return [(-1, any_, any_, L.GF_ALL)]
    if method == 1:
        return [(a, any_, any_, L.GF_ALL) for a in range(3)]
    out = []
    for send, recv, kind in ((-1, 1, L.GF_F2C), (0, 0, L.GF_FLAT), (1, -1, L.GF_C2F)):
        out += [(a, send, recv, kind) for a in range(3)]
    return out


def split_by_call(c, rank_of, dtype):
    """Each rank's messages cut into exchange calls by the send plans."""
    g = c["mesh"]
    out = []
    for r in range(int(c["ranks"])):
        seq, pos = [], 0
        for axis, sl, _, _ in calls(int(c["method"])):
            p = CommPlan(g["blocks"], g["n"], g["vc"], dtype, rank_of, r, L.COMM_SEND, axis, sl)
            seq.append(c["msgs"][r][pos:pos + len(p.peers)])
            pos += len(p.peers)
        assert pos == len(c["msgs"][r]), (r, pos, len(c["msgs"][r]))
        out.append(seq)
    return out


def test_partition_is_the_references():
    assert partition(10, 3).tolist() == [0, 0, 0, 0, 1, 1, 1, 2, 2, 2]
    assert partition(4, 4).tolist() == [0, 1, 2, 3]
    assert np.bincount(partition(3816, 8)).tolist() == [477] * 8


@pytest.mark.parametrize("name", COMM)
def test_plans_name_the_references_peers_and_message_sizes(name):
    c = load_comm(name)
    g = c["mesh"]
    ranks = int(c["ranks"])
    dtype = np.float64 if str(c["dtype"]) == "d" else np.float32
    rank_of = partition(len(g["blocks"]), ranks)
    assert [int(v) for v in c["start"]] == [int(np.searchsorted(rank_of, r)) for r in range(ranks)]
    per_call = split_by_call(c, rank_of, dtype)
    for ci, (axis, sl, rl, _) in enumerate(calls(int(c["method"]))):
        for r in range(ranks):
            send = CommPlan(g["blocks"], g["n"], g["vc"], dtype, rank_of, r, L.COMM_SEND, axis, sl)
            assert [(p[0], p[4]) for p in send.peers] == [(d, len(b)) for d, b in per_call[r][ci]], (ci, r)
            # The receive plan expects, from each peer, what that peer sent.
            recv = CommPlan(g["blocks"], g["n"], g["vc"], dtype, rank_of, r, L.COMM_RECV, axis, rl)
            got = [(s, len(b)) for s in range(ranks) for d, b in per_call[s][ci] if d == r]
            assert [(p[0], p[4]) for p in recv.peers] == got, (ci, r)
            for plan in (send, recv):
                assert plan.total_bytes == sum(p[4] for p in plan.peers)
                for rank_, first, nf, off, nbytes in plan.peers:
                    assert off % 8 == 0 and nbytes % 8 == 0
                    faces = plan.faces[first:first + nf]
                    assert all(rank_of[f.block] == r for f in faces)
                    assert all(axis < 0 or f.face // 2 == axis for f in faces)


@pytest.mark.parametrize("name,split", [("comm5_periodic_x_d", "partition"), ("comm5_periodic_x_d", "scattered"),
                                        ("comm4_sphere_vc2_leveldiff_f", "scattered")])
def test_plans_cover_every_link_once(name, split):
    """Over all ranks, the send and receive plans of an AtOnce exchange
    list each link between ranks exactly once, the receiver's faces
    matching the sender's, slot by slot: for the reference's contiguous
    partition and for blocks scattered over ranks at random."""
    c = load_comm(name)
    g = c["mesh"]
    blocks = g["blocks"]
    ranks = int(c["ranks"])
    if split == "partition":
        rank_of = partition(len(blocks), ranks)
    else:
        rank_of = np.random.default_rng(ranks).integers(0, ranks, len(blocks)).astype(np.int32)
    sends = [CommPlan(blocks, g["n"], g["vc"], np.float64, rank_of, r, L.COMM_SEND) for r in range(ranks)]
    recvs = [CommPlan(blocks, g["n"], g["vc"], np.float64, rank_of, r, L.COMM_RECV) for r in range(ranks)]
    links = 0
    for r in range(ranks):
        for peer, first, nf, _, nbytes in sends[r].peers:
            back = [p for p in recvs[peer].peers if p[0] == r]
            assert len(back) == 1 and back[0][2] == nf and back[0][4] == nbytes
            rf = recvs[peer].faces[back[0][1]:back[0][1] + nf]
            for sf, qf in zip(sends[r].faces[first:first + nf], rf):
                # The slot the sender fills is the one the receiver reads:
                # its block is the sender's neighbour across that face.
                assert qf.block in [int(v) for v in blocks["nb_id"][sf.block][sf.face]]
                assert sf.block in [int(v) for v in blocks["nb_id"][qf.block][qf.face]]
                assert sf.face == qf.face ^ 1
                assert sf.offset - sends[r].faces[first].offset == qf.offset - rf[0].offset
            links += nf
    cross = 0
    for b in range(len(blocks)):
        for f in range(6):
            ld = int(blocks["nb_level_diff"][b][f])
            for s in range(4 if ld == 1 else 1):
                nb = int(blocks["nb_id"][b][f][s])
                cross += nb >= 0 and rank_of[nb] != rank_of[b]
    assert links == cross > 0


def test_plan_padding_and_argument_checks():
    """FP32 faces of n = 2 blocks: 4-byte F2C slots take 8 bytes."""
    c = load_comm("comm3_simple_n2_f")
    g = c["mesh"]
    blocks = g["blocks"]
    rank_of = partition(len(blocks), 3)
    seen = 0
    for r in range(3):
        for d in (L.COMM_SEND, L.COMM_RECV):
            p = CommPlan(blocks, 2, 1, np.float32, rank_of, r, d)
            slots = [(f.offset, f.kind) for f in p.faces[:p.nfaces]]
            seen += sum(k == L.COMM_F2C for _, k in slots)
            for (o0, k0), (o1, _) in zip(slots, slots[1:]):
                assert (o1 - o0) * 4 == (8 if k0 == L.COMM_F2C else 16)
    assert seen > 0
    lib = L.load()
    br = np.ascontiguousarray(rank_of, dtype=np.int32)
    n1, n2, n3 = (L.ctypes.c_int32(), L.ctypes.c_int64(), L.ctypes.c_int64())

    def plan(n=2, vc=1, dtype=L.F32, rank=0, direction=L.COMM_SEND, axis=-1, level=L.COMM_ANY_LEVEL, ranks=br):
        return lib.bcm_comm_plan(blocks.ctypes.data, len(blocks), n, vc, dtype, ranks.ctypes.data, rank, direction,
                                 axis, level, None, 0, L.ctypes.byref(n1), None, 0, L.ctypes.byref(n2),
                                 L.ctypes.byref(n3))

    assert plan() == L.OK
    for bad in (dict(n=3), dict(vc=2), dict(dtype=5), dict(rank=-1), dict(direction=2), dict(axis=3), dict(level=-2),
                dict(ranks=np.full(len(blocks), -1, np.int32))):
        assert plan(**bad) == L.ERR_ARG, bad
    # Arrays too small for the plan.
    peers = (L.CommPeer * 1)()
    faces = (L.CommFace * 1)()
    assert lib.bcm_comm_plan(blocks.ctypes.data, len(blocks), 2, 1, L.F32, br.ctypes.data, 1, L.COMM_SEND, -1,
                             L.COMM_ANY_LEVEL, peers, 1, L.ctypes.byref(n1), faces, 1, L.ctypes.byref(n2),
                             L.ctypes.byref(n3)) == L.ERR_ARG
    # One rank: nothing leaves it.
    one = CommPlan(blocks, 2, 1, np.float32, np.zeros(len(blocks), np.int32), 0, L.COMM_SEND)
    assert one.peers == [] and one.total_bytes == 0
