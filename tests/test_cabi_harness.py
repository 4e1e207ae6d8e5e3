"""The go-ethereum C-ABI call sequence replayed by a C++ caller.

tests/cpp/geth_sequence.cpp links libgmpt_amd.so through include/ only (the
way a cgo or C++ caller would) and runs preprocess -> build_mpt_2phase /
build_mpt_olc for the state, transaction and receipt tries over several
blocks -> get_all_nodes -> init_mpt / insert_mpt_* / free_mpt, then the
GpuMPT::Compress::MPT class (mpt-with-compress/lib/libgmpt.cu:25-331,
gpu_mpt.cuh:69-152).  Every root and exported node is checked against the
oracle.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
EXE = os.path.join(CPP, "geth_sequence")


def _blocks(workloads, nblocks=4, seed=60):
    """nblocks x 3 tries; later blocks overwrite some earlier state keys."""
    rng = np.random.default_rng(seed)
    out = []
    prev_keys = None
    for blk in range(nblocks):
        row = []
        for t in range(3):
            n = int(rng.integers(200, 900))
            keys = workloads.unique_random_keys(rng, n, 32)
            if t == 0 and prev_keys is not None:
                m = min(n, len(prev_keys)) // 4
                keys[:m] = prev_keys[:m]  # updates of existing state keys
            flat, lens = workloads.random_values(rng, n, 30, 120)
            row.append(workloads.batch_from_fixed(keys, flat, lens))
            if t == 0:
                prev_keys = keys
        out.append(row)
    return out


def _write(path, blocks):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(blocks)))
        for row in blocks:
            for b in row:
                f.write(struct.pack("<IQQ", b.n, b.keys_hexs.size, b.values_bytes.size))
                f.write(b.keys_hexs.tobytes())
                f.write(b.keys_indexs.astype("<i4").tobytes())
                f.write(b.values_bytes.tobytes())
                f.write(b.values_indexs.astype("<i8").tobytes())


def test_harness_builds():
    r = subprocess.run(["make", "-s", "-C", CPP], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(EXE)


@pytest.mark.gpu
def test_geth_call_sequence(workloads, oracle, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run make -C tests/cpp)")
    blocks = _blocks(workloads)
    path = str(tmp_path / "batches.bin")
    _write(path, blocks)
    r = subprocess.run([EXE, path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines()]

    # cumulative oracle tries, one per TrieType
    tries = [oracle.OracleMPT() for _ in range(3)]
    want_build = {}
    for blk, row in enumerate(blocks):
        for t, b in enumerate(row):
            tries[t].puts(b)
            tries[t].hash()
            want_build[(blk, t)] = tries[t].root().hex()
    got_build = {(int(l[1]), int(l[2])): l[3] for l in lines if l[0] == "build"}
    assert got_build == want_build

    # get_all_nodes: every node on the last block's state-key paths
    state = tries[0]
    all_nodes = {h + bytes(32 - len(h)): e for h, e in state.export()}
    got_nodes = [(bytes.fromhex(l[2]), bytes.fromhex(l[3])) for l in lines if l[0] == "nodes"]
    assert len(got_nodes) == int([l for l in lines if l[0] == "nodes"][0][1])
    assert len({h for h, _ in got_nodes}) == len(got_nodes)
    for h, e in got_nodes:
        assert all_nodes.get(h) == e
    root = state.root()
    assert root + bytes(32 - len(root)) in {h for h, _ in got_nodes}

    # init_mpt / insert_mpt_*: state batches only, cumulative
    o = oracle.OracleMPT()
    for blk, row in enumerate(blocks):
        o.puts(row[0])
        o.hash()
        got = [l[2] for l in lines if l[0] == "insert" and int(l[1]) == blk]
        assert got == [o.root().hex()], blk

    assert [l[1] for l in lines if l[0] == "class"] == [o.root().hex()]
    assert [l[1] for l in lines if l[0] == "gets"] == ["1"]
