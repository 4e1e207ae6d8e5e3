"""GPU parity: PhaseNU / LockNU / PhaseHC / get / node export through the C-ABI,
compared bit-for-bit with the CPU oracle on the same seeded inputs.

Model: the reference's relative CPU==GPU root checks (mpt-with-compress/test/
trie.cu:1451, 4093), its get tests (trie.cu:153-197) and override test
(trie.cu:293-328, where this build fixes last-writer-wins on the GPU too).
"""
import ctypes
import json
import os
import random

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu

MODES = ["2phase", "latching"]


def _put(t, b, mode, hps=None):
    fn = t.puts_2phase_with_valuehp if mode == "2phase" else t.puts_latching_with_valuehp_v2
    targets, n = fn(b, hps)
    t.hash_onepass_v2(targets, n)


def _build(gmpt, b, mode):
    t = gmpt.MPT(device=0)
    _put(t, b, mode)
    return t


def _rand_batch(workloads, seed, n, vlo, vhi, key_len=32):
    rng = np.random.default_rng(seed)
    keys = workloads.unique_random_keys(rng, n, key_len)
    flat, lens = workloads.random_values(rng, n, vlo, vhi)
    return workloads.batch_from_fixed(keys, flat, lens)


@pytest.mark.parametrize("kmode", [0, 1], ids=["lane", "spread"])
def test_keccak_matches_oracle(gmpt, oracle, kmode):
    with open(os.path.join(GOLDEN, "keccak256.json")) as f:
        vecs = json.load(f)
    msgs = [bytes.fromhex(v["msg"]) for v in vecs]
    got = gmpt.keccak256_batch(msgs, kmode)
    assert [g.hex() for g in got] == [v["digest"] for v in vecs]
    rng = random.Random(3)
    msgs = [bytes(rng.getrandbits(8) for _ in range(L)) for L in list(range(0, 420)) + [543, 544, 1087, 1088, 2000]]
    got = gmpt.keccak256_batch(msgs, kmode)
    for m, g in zip(msgs, got):
        assert g == oracle.keccak256(m), len(m)


@pytest.mark.parametrize("mode", MODES)
def test_doe_dog_dogglesworth(gmpt, workloads, mode):
    with open(os.path.join(GOLDEN, "roots.json")) as f:
        g = json.load(f)
    b = workloads.batch_from_bytes([b"doe", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    t = _build(gmpt, b, mode)
    assert t.get_root_hash().hex() == g["doe_dog_dogglesworth"]
    assert t.get_values(b) == [b"reindeer", b"puppy", b"cat"]
    miss = workloads.batch_from_bytes([b"do", b"dogg", b"cat"], [b"", b"", b""])
    assert t.get_values(miss) == [None, None, None]


@pytest.mark.parametrize("mode", MODES)
def test_gets_return_host_pointers(gmpt, workloads, mode):
    # gets_parallel returns the host pointer handed over at insert
    # (gpu_mpt_kernels.cuh:2285-2290); read the bytes back through it.
    b = workloads.batch_from_bytes([b"doe", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    base = b.values_bytes.ctypes.data
    hps = np.array([base + int(b.values_indexs[2 * i]) for i in range(b.n)], np.uint64)
    t = gmpt.MPT(device=0)
    _put(t, b, mode, hps)
    ptrs, sizes = t.gets_parallel(b)
    for i in range(b.n):
        assert ctypes.string_at(int(ptrs[i]), int(sizes[i])) == b.value(i)


@pytest.mark.parametrize("mode", MODES)
def test_override_last_writer_wins(gmpt, workloads, mode):
    b = workloads.batch_from_bytes([b"dog", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    t = _build(gmpt, b, mode)
    assert t.get_values(b) == [b"puppy", b"puppy", b"cat"]


@pytest.mark.parametrize("mode", MODES)
def test_config1_golden_root(gmpt, workloads, mode):
    with open(os.path.join(GOLDEN, "roots.json")) as f:
        g = json.load(f)
    t = _build(gmpt, workloads.config1_oracle(), mode)
    assert t.get_root_hash().hex() == g["config1"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,vlo,vhi", [(1, 40, 40), (2, 1, 3), (3, 33, 300), (17, 1, 40), (1000, 0, 8),
                                       (5000, 70, 110), (3000, 130, 560), (20000, 1, 200)])
def test_random_vs_oracle(gmpt, workloads, oracle, mode, n, vlo, vhi):
    b = _rand_batch(workloads, 100 + n, n, vlo, vhi)
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)


@pytest.mark.parametrize("mode", MODES)
def test_variable_length_keys(gmpt, workloads, oracle, mode):
    rng = random.Random(17)
    keys = set()
    while len(keys) < 1500:
        keys.add(bytes(rng.choice([0, 1, 2, 0x10, 0xff]) for _ in range(rng.randint(1, 6))))
    keys = list(keys)
    rng.shuffle(keys)
    vals = [bytes([rng.getrandbits(8)]) * rng.randint(1, 70) for _ in keys]
    b = workloads.batch_from_bytes(keys, vals)
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)
    assert t.get_values(b) == vals


@pytest.mark.parametrize("mode", MODES)
def test_duplicates_in_batch(gmpt, workloads, oracle, mode):
    rng = np.random.default_rng(9)
    keys = workloads.unique_random_keys(rng, 400, 32)
    idx = rng.integers(0, 400, size=3000)
    flat, lens = workloads.random_values(rng, 3000, 30, 90)
    b = workloads.batch_from_fixed(keys[idx], flat, lens)
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)
    last = {}
    for i in range(b.n):
        last[b.key(i)] = b.value(i)
    assert t.get_values(b) == [last[b.key(i)] for i in range(b.n)]


@pytest.mark.parametrize("modes", [("2phase", "2phase"), ("latching", "latching"), ("2phase", "latching"),
                                   ("latching", "2phase")])
def test_incremental_batches(gmpt, workloads, oracle, modes):
    rng = np.random.default_rng(5)
    base = _rand_batch(workloads, 5, 6000, 30, 120)
    # second batch: 1500 overwrites of existing keys + 2500 new keys, shuffled
    upd = base.subset(rng.choice(6000, 1500, replace=False))
    fresh = _rand_batch(workloads, 55, 2500, 1, 200)
    flat, lens = workloads.random_values(rng, 1500, 40, 80)
    upd = workloads.batch_from_hex([upd.key(i) for i in range(upd.n)],
                                   [flat[int(lens[:i].sum()):int(lens[:i + 1].sum())].tobytes() for i in range(1500)])
    second = workloads.concat_batches([upd, fresh])
    second = second.subset(rng.permutation(second.n))
    t = gmpt.MPT(device=0)
    _put(t, base, modes[0])
    o = oracle.OracleMPT()
    o.puts(base)
    o.hash()
    assert t.get_root_hash() == o.root()
    _put(t, second, modes[1])
    o.puts(second)
    o.hash()
    assert t.get_root_hash() == o.root()
    sample = second.subset(range(0, second.n, 37))
    assert t.get_values(sample) == o.gets(sample)


@pytest.mark.parametrize("modes", [("2phase", "2phase"), ("latching", "latching"), ("2phase", "latching")])
def test_trie_state_in_host_memory(gmpt, workloads, oracle, modes):
    """Out-of-HBM tries (the reference's UMMPT, um_mpt.cuh:13-237): with
    pool_memory = 1 the node pools, key and value arenas and dirty list live
    in pinned host memory the kernels reach over the host link.  A build, a
    second batch of overwrites and new keys (descent marks, splits, latches
    on host memory), gets and the node export all match the oracle and a
    device-pool forest."""
    rng = np.random.default_rng(71)
    base = _rand_batch(workloads, 71, 20000, 30, 120)
    upd = base.subset(rng.choice(base.n, 3000, replace=False))
    flat, lens = workloads.random_values(rng, 3000, 40, 80)
    upd = workloads.batch_from_hex([upd.key(i) for i in range(upd.n)],
                                   [flat[int(lens[:i].sum()):int(lens[:i + 1].sum())].tobytes() for i in range(3000)])
    second = workloads.concat_batches([upd, _rand_batch(workloads, 72, 5000, 1, 200)])
    second = second.subset(rng.permutation(second.n))
    t = gmpt.MPT(device=0)
    t.set_option("pool_memory", 1)
    assert t.get_option("pool_memory") == 1
    d = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    for b, mode in ((base, modes[0]), (second, modes[1])):
        _put(t, b, mode)
        _put(d, b, mode)
        o.puts(b)
        o.hash()
        assert t.get_root_hash() == o.root()
    sample = second.subset(range(0, second.n, 41))
    assert t.get_values(sample) == o.gets(sample)
    # LockNU numbers the nodes its latched splits create in arrival order, so
    # the two forests list the same nodes in different orders
    assert sorted(t.flush_dirty_nodes(sample)) == sorted(d.flush_dirty_nodes(sample))
    assert t.stats()["pool_bytes_reserved"] > 0
    t.close()
    d.close()


def test_repeated_updates_rewrite_values_in_place(gmpt, workloads, oracle):
    """PhaseNU update-only batches rewrite the keys' value nodes in place and
    keep no packed keys: across five rounds of updates the pools grow by the
    staged value bytes only, and every root and value matches the oracle."""
    rng = np.random.default_rng(61)
    base = _rand_batch(workloads, 61, 4000, 30, 90)
    t = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    _put(t, base, "2phase")
    o.puts(base)
    o.hash()
    assert t.get_root_hash() == o.root()
    used0 = t.stats()["pool_bytes_used"]
    staged = 0
    for r in range(5):
        flat, lens = workloads.random_values(rng, base.n, 20, 140)
        upd = workloads.batch_from_hex([base.key(i) for i in range(base.n)],
                                       [flat[int(lens[:i].sum()):int(lens[:i + 1].sum())].tobytes()
                                        for i in range(base.n)])
        upd = upd.subset(rng.permutation(upd.n))
        _put(t, upd, "2phase")
        o.puts(upd)
        o.hash()
        assert t.get_root_hash() == o.root(), r
        staged += int(lens.sum()) + 8 + 256  # one staged value buffer per batch (8-byte slack, 256-byte aligned)
    grown = t.stats()["pool_bytes_used"] - used0
    assert grown <= staged, (grown, staged)
    sample = base.subset(range(0, base.n, 13))
    assert t.get_values(sample) == o.gets(sample)
    # new keys after the updates still land next to the kept keys
    fresh = _rand_batch(workloads, 62, 3000, 30, 90)
    _put(t, fresh, "2phase")
    o.puts(fresh)
    o.hash()
    assert t.get_root_hash() == o.root()
    assert t.get_values(fresh) == o.gets(fresh)


@pytest.mark.parametrize("mode", MODES)
def test_many_small_batches_split_shorts(gmpt, workloads, oracle, mode):
    # keys sharing long prefixes split existing short nodes at every depth
    rng = random.Random(23)
    keys = sorted({bytes([0xAB, 0xCD]) + bytes(rng.choice([0, 1, 0x11]) for _ in range(rng.randint(1, 8)))
                   for _ in range(600)})
    rng.shuffle(keys)
    vals = [bytes([i % 251]) * (33 + i % 40) for i in range(len(keys))]
    t = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    for lo in range(0, len(keys), 50):
        b = workloads.batch_from_bytes(keys[lo:lo + 50], vals[lo:lo + 50])
        _put(t, b, mode)
        o.puts(b)
        o.hash()
        assert t.get_root_hash() == o.root(), lo


@pytest.mark.parametrize("mode", MODES)
def test_node_export_matches_oracle(gmpt, workloads, oracle, mode):
    b = _rand_batch(workloads, 77, 3000, 20, 150)
    t = _build(gmpt, b, mode)
    got = t.flush_dirty_nodes(b)
    o = oracle.OracleMPT()
    o.puts(b)
    o.hash()
    want = [(h + bytes(32 - len(h)), e) for h, e in o.export()]
    assert sorted(got) == sorted(want)


@pytest.mark.parametrize("mode", MODES)
def test_multi_trie_batch(gmpt, workloads, oracle, mode):
    rng = np.random.default_rng(8)
    batches = [_rand_batch(workloads, 800 + i, int(s), 30, 100) for i, s in enumerate(rng.integers(1, 700, 8))]
    allb = workloads.concat_batches(batches, trie_ids=list(range(8)))
    allb = allb.subset(rng.permutation(allb.n))
    t = gmpt.MPT(device=0, num_tries=8)
    _put(t, allb, mode)
    for i, b in enumerate(batches):
        assert t.get_root_hash(i) == oracle.root_of(b), i


def test_device_resident_batch(gmpt, workloads, oracle):
    b = _rand_batch(workloads, 31, 4000, 70, 110)
    db = gmpt.DeviceBatch(b, adopt_values=True)
    t = gmpt.MPT(device=0)
    targets, n = t.puts_2phase_with_valuehp(db)
    t.hash_onepass_v2(targets, n)
    assert t.get_root_hash() == oracle.root_of(b)
    assert t.get_values(b.subset(range(0, 4000, 101))) == [b.value(i) for i in range(0, 4000, 101)]
    t.close()
    db.free()


@pytest.mark.parametrize("mode", MODES)
def test_malformed_key_is_rejected(gmpt, workloads, oracle, mode):
    t = gmpt.MPT(device=0)
    bad = workloads.batch_from_hex([bytes([1, 2, 16]), bytes([1, 17, 16])], [b"x" * 40, b"y" * 40])
    with pytest.raises(gmpt.GmptError):
        _put(t, bad, mode)
    no_term = workloads.batch_from_hex([bytes([1, 2, 3])], [b"z" * 40])
    with pytest.raises(gmpt.GmptError):
        _put(t, no_term, mode)
    good = _rand_batch(workloads, 3, 100, 40, 60)
    _put(t, good, mode)
    assert t.get_root_hash() == oracle.root_of(good)


def test_reference_c_abi(gmpt, workloads, oracle):
    L = gmpt.gmpt.lib()
    b = _rand_batch(workloads, 44, 2000, 40, 90)
    args = [b.keys_hexs.ctypes.data, b.keys_indexs.ctypes.data, b.values_bytes.ctypes.data,
            b.values_indexs.ctypes.data, None, b.n]
    assert L.preprocess() == 1
    want = oracle.root_of(b)
    h = L.build_mpt_2phase(0, *args)
    assert ctypes.string_at(h, 32) == want
    h2 = L.build_mpt_olc(2, *args)
    assert ctypes.string_at(h2, 32) == want
    ns = L.get_all_nodes(0, b.keys_hexs.ctypes.data, b.keys_indexs.ctypes.data, b.n)
    assert ns.num > b.n
    m = L.init_mpt()
    h3 = L.insert_mpt_2phase(m, *args)
    assert ctypes.string_at(h3, 32) == want
    h4 = L.insert_mpt_olc(m, *args)
    assert ctypes.string_at(h4, 32) == want
    L.free_mpt(m)
    for p in (h, h2, h3, h4, ns.hashs, ns.encs, ns.encs_indexs):
        L.gmpt_free_host(p)


@pytest.mark.parametrize("spread_max", [0, 1 << 30], ids=["lane-only", "spread-only"])
def test_keccak_layouts_give_same_root(gmpt, workloads, oracle, spread_max):
    b = _rand_batch(workloads, 91, 20000, 20, 300)
    t = gmpt.MPT(device=0)
    t.set_option("hc_spread_max", spread_max)
    assert t.get_option("hc_spread_max") == spread_max
    targets, n = t.puts_2phase_with_valuehp(b)
    t.hash_onepass_v2(targets, n)
    assert t.get_root_hash() == oracle.root_of(b)


@pytest.mark.parametrize("mode", MODES)
def test_large_host_batch_through_pinned_staging(gmpt, workloads, oracle, mode):
    # a host batch large enough (index pairs >= 8 MB) for the chunked pinned
    # stager: every array crosses several 8 MB chunks of the staging ring
    b = _rand_batch(workloads, 97, 400000, 70, 110)
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)
    sub = b.subset(range(0, b.n, 4099))
    assert t.get_values(sub) == [sub.value(i) for i in range(sub.n)]
    t.close()


@pytest.mark.parametrize("fuse_inner", [0, 1], ids=["encode+permute", "lds-rows"])
@pytest.mark.parametrize("mode", MODES)
def test_inner_level_kernels_give_same_root(gmpt, workloads, oracle, fuse_inner, mode):
    # every inner level above the spread threshold (0: all of them) by the
    # LDS-row kernels or the 16-lane encoder + scratch + permutation; tiny
    # values make embedded children and short nodes over them (the generic
    # encoder inside the row kernels), prefix keys put values in slot 16
    cases = [_rand_batch(workloads, 92, 30000, 20, 300), _rand_batch(workloads, 96, 4000, 0, 3)]
    rng = np.random.default_rng(98)
    ks = [bytes(rng.integers(0, 256, int(rng.integers(1, 6)), dtype=np.uint8)) for _ in range(3000)]
    ks = list(dict.fromkeys(ks))
    cases.append(workloads.batch_from_bytes(ks, [bytes(rng.integers(0, 256, 40, dtype=np.uint8)) for _ in ks]))
    for b in cases:
        t = gmpt.MPT(device=0)
        t.set_option("hc_spread_max", 0)
        t.set_option("hc_fuse_inner", fuse_inner)
        assert t.get_option("hc_fuse_inner") == fuse_inner
        _put(t, b, mode)
        assert t.get_root_hash() == oracle.root_of(b)
        t.close()


def _leaf_case(workloads, case):
    if case == "c2":
        return _rand_batch(workloads, 93, 5000, 70, 110)
    if case == "block-edge":  # 129 leaves: a partial second block of the fused kernel
        return _rand_batch(workloads, 94, 129, 70, 110)
    if case == "tiny-values":  # raw single-byte and empty values
        return _rand_batch(workloads, 95, 3000, 0, 2)
    if case == "long-keys":  # key remainders over 64 nibbles take the generic encoder
        return _rand_batch(workloads, 96, 2000, 10, 60, key_len=44)
    if case == "short-keys":  # leaf encodings under 32 bytes keep the raw encoding
        rng = random.Random(97)
        keys = list({bytes(rng.getrandbits(8) for _ in range(rng.randint(1, 3))) for _ in range(3000)})
        vals = [bytes(rng.getrandbits(8) for _ in range(rng.randint(0, 40))) for _ in keys]
        return workloads.batch_from_bytes(keys, vals)
    return _rand_batch(workloads, 98, 8000, 20, 600)  # mixed: one to five blocks per leaf


@pytest.mark.parametrize("fuse", [0, 1], ids=["leaf-kernels", "leaf-fused"])
@pytest.mark.parametrize("case", ["c2", "block-edge", "tiny-values", "long-keys", "short-keys", "mixed"])
def test_leaf_paths_match_oracle(gmpt, workloads, oracle, fuse, case):
    """The fused leaf kernel (encoding through LDS) and the encode + scratch +
    permutation kernels give the oracle's root and the same per-pass counts."""
    b = _leaf_case(workloads, case)
    t = gmpt.MPT(device=0)
    t.set_option("hc_fuse_leaves", fuse)
    assert t.get_option("hc_fuse_leaves") == fuse
    targets, n = t.puts_2phase_with_valuehp(b)
    t.hash_onepass_v2(targets, n)
    assert t.get_root_hash() == oracle.root_of(b)
    st = t.stats()
    t2 = gmpt.MPT(device=0)
    t2.set_option("hc_fuse_leaves", 1 - fuse)
    targets, n = t2.puts_2phase_with_valuehp(b)
    t2.hash_onepass_v2(targets, n)
    st2 = t2.stats()
    for k in ("perms", "enc_bytes", "leaf_perms", "leaf_bytes", "leaf_nodes", "hashed_nodes"):
        assert st[k] == st2[k], k
    with pytest.raises(gmpt.GmptError):
        t.set_option("hc_fuse_leaves", 2)


def test_autotune_keeps_results(gmpt, workloads, oracle):
    b = _rand_batch(workloads, 92, 6000, 70, 110)
    t = gmpt.MPT(device=0)
    v, ms = t.autotune(b)
    assert t.get_option("hc_spread_max") == v and ms > 0
    with pytest.raises(gmpt.GmptError):
        t.set_option("no_such_option", 1)
    targets, n = t.puts_2phase_with_valuehp(b)
    t.hash_onepass_v2(targets, n)
    assert t.get_root_hash() == oracle.root_of(b)


@pytest.mark.parametrize("mode", MODES)
def test_mixed_read_write(gmpt, workloads, oracle, mode):
    """gets + puts in one batch (the *_with_read entry points): reads see the
    trie before the batch's writes, in batch order; writes land as usual."""
    rng = np.random.default_rng(123)
    pre = _rand_batch(workloads, 70, 3000, 30, 100)
    pre_hps = np.array([pre.values_bytes.ctypes.data + int(pre.values_indexs[2 * i]) for i in range(pre.n)],
                       np.uint64)
    t = gmpt.MPT(device=0)
    _put(t, pre, mode, pre_hps)
    # ops: reads of existing keys, reads of absent keys, updates, inserts,
    # and keys both read and written in the batch
    fresh = _rand_batch(workloads, 71, 800, 30, 100)
    picks = rng.choice(pre.n, 1200, replace=False)
    ops_keys, ops_vals, flags, want_reads = [], [], [], []
    for j, i in enumerate(picks):
        ops_keys.append(bytes(pre.key(int(i))))
        if j % 3 == 0:  # update
            ops_vals.append(bytes(rng.integers(0, 256, 50, dtype=np.uint8)))
            flags.append(1)
        else:
            ops_vals.append(b"")
            flags.append(0)
            want_reads.append((int(pre_hps[i]), pre.value(int(i)).__len__()))
    for i in range(fresh.n):
        if i % 2 == 0:  # insert
            ops_keys.append(bytes(fresh.key(i)))
            ops_vals.append(fresh.value(i))
            flags.append(1)
        else:  # read of an absent key
            ops_keys.append(bytes(fresh.key(i)))
            ops_vals.append(b"")
            flags.append(0)
            want_reads.append((0, 0))
    ops = workloads.batch_from_hex(ops_keys, ops_vals)
    targets, n, reads = t.puts_with_read(ops, np.array(flags, np.uint8), 0 if mode == "2phase" else 1)
    t.hash_onepass_v2(targets, n)
    assert reads == want_reads
    o = oracle.OracleMPT()
    o.puts(pre)
    w = [i for i in range(ops.n) if flags[i] == 1]
    o.puts(ops.subset(w))
    o.hash()
    assert t.get_root_hash() == o.root()
    assert t.get_values(ops.subset(w)) == [ops.value(i) for i in w]
    with pytest.raises(gmpt.GmptError):
        t.puts_with_read(ops, np.full(ops.n, 2, np.uint8))


def test_long_runs_of_equal_leading_window_bits(gmpt, workloads, oracle):
    """Keys whose sort windows share their leading 24 bits in runs far longer
    than the sort's fix-up handles (the 64-bit re-sort path), plus keys equal
    in the whole 16-nibble window (the full-key path)."""
    rng = np.random.default_rng(33)
    n = 3000
    keys = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    keys[:, 0] = rng.integers(0, 2, n) * 0x10  # first nibble 0 or 1, so the common prefix is empty
    keys[:, 1:3] = 0x5a                           # the next 5 nibbles are fixed
    keys[: n // 3, 3:8] = 0x77                    # a third also share nibbles 6..15
    keys = np.unique(keys, axis=0)
    flat, lens = workloads.random_values(rng, len(keys), 40, 90)
    b = workloads.batch_from_fixed(keys, flat, lens)
    b = b.subset(rng.permutation(b.n))
    t = gmpt.MPT(device=0)
    _put(t, b, "2phase")
    assert t.get_root_hash() == oracle.root_of(b)
    assert t.stats()["fallback_sorts"] == 1


@pytest.mark.parametrize("mode", MODES)
def test_long_runs_in_several_tries(gmpt, workloads, oracle, mode):
    """The re-sort by the window's varying bits (here 1 + 40 of 64) and the
    whole-key sort when the batch spans tries (the trie pass follows the
    window sort), with repeated keys in each trie."""
    rng = np.random.default_rng(34)
    batches = []
    for tr in range(3):
        n = 1500
        keys = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        keys[:, 0] = rng.integers(0, 2, n) * 0x10
        keys[:, 1:3] = 0x5a
        keys[: n // 3, 3:8] = 0x77
        keys = np.unique(keys, axis=0)
        keys = np.concatenate([keys, keys[rng.integers(0, len(keys), 200)]])  # repeats: last write wins
        flat, lens = workloads.random_values(rng, len(keys), 40, 90)
        batches.append(workloads.batch_from_fixed(keys, flat, lens))
    allb = workloads.concat_batches(batches, trie_ids=[0, 1, 2])
    allb = allb.subset(rng.permutation(allb.n))
    t = gmpt.MPT(device=0, num_tries=3)
    _put(t, allb, mode)
    for i in range(3):
        sub = allb.subset(np.nonzero(allb.trie_ids == i)[0])
        assert t.get_root_hash(i) == oracle.root_of(sub), i


@pytest.mark.parametrize("mode", MODES)
def test_long_values(gmpt, workloads, oracle, mode):
    """Leaves of many Keccak blocks (values up to 3,000 bytes) and a few
    single-byte and empty values."""
    rng = np.random.default_rng(44)
    b = _rand_batch(workloads, 45, 1500, 900, 3000)
    small = _rand_batch(workloads, 46, 40, 0, 1)
    allb = workloads.concat_batches([b, small], trie_ids=[0, 0])
    allb.trie_ids = None
    allb = allb.subset(rng.permutation(allb.n))
    t = gmpt.MPT(device=0)
    _put(t, allb, mode)
    assert t.get_root_hash() == oracle.root_of(allb)


@pytest.mark.parametrize("mode", MODES)
def test_skewed_user_keys_on_preloaded_trie(gmpt, workloads, oracle, mode):
    """The skewed-update shape at reduced size: 'user' + 19-digit keys (a
    31-nibble shared prefix, then digit branches under one-nibble extensions),
    a preloaded trie, then Zipf(0.99) overwrites and inserts with many
    repeated keys, so concurrent inserts meet on shared branches."""
    pre, ops = workloads.config3_lock(seed=7, n_preload=150_000, n_ops=100_000, universe=400_000, value_len=40)
    t = gmpt.MPT(device=0)
    _put(t, pre, "2phase")
    o = oracle.OracleMPT()
    o.puts(pre)
    o.hash()
    assert t.get_root_hash() == o.root()
    _put(t, ops, mode)
    assert t.stats()["ext_direct_nodes"] >= 16384  # one-nibble extensions hashed as leaf-shaped nodes
    o.puts(ops)
    o.hash()
    assert t.get_root_hash() == o.root()
    sample = ops.subset(range(0, ops.n, 97))
    assert t.get_values(sample) == o.gets(sample)


@pytest.mark.parametrize("seed", [61, 62, 63])
def test_latched_inserts_split_fresh_nodes(gmpt, workloads, oracle, seed):
    """Keys built so that concurrent inserts of one wavefront split the same
    fresh short nodes at many depths: clusters of keys sharing prefixes of
    random lengths, inserted with LockNU into an empty trie and then into the
    result, checked against the oracle after each batch."""
    rng = np.random.default_rng(seed)
    keys = []
    for _ in range(300):
        base = rng.integers(0, 256, 32, dtype=np.uint8)
        for _ in range(int(rng.integers(2, 40))):
            k = base.copy()
            cut = int(rng.integers(1, 32))
            k[cut:] = rng.integers(0, 256, 32 - cut, dtype=np.uint8)
            keys.append(k)
    keys = np.unique(np.array(keys), axis=0)
    keys = keys[rng.permutation(len(keys))]
    half = len(keys) // 2
    t = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    for part in (keys[:half], keys[half:]):
        flat, lens = workloads.random_values(rng, len(part), 1, 120)
        b = workloads.batch_from_fixed(part, flat, lens)
        _put(t, b, "latching")
        o.puts(b)
        o.hash()
        assert t.get_root_hash() == o.root()
    sample = workloads.batch_from_fixed(keys[::11], *workloads.random_values(rng, len(keys[::11]), 1, 2))
    assert t.get_values(sample) == o.gets(sample)


@pytest.mark.parametrize("modes", [("2phase", "latching"), ("latching", "latching"), ("latching", "2phase")])
def test_prefix_keys_incremental(gmpt, workloads, oracle, modes):
    """Keys that are prefixes of other keys (values in branch slot 16): a
    second batch overwrites some of them, adds new prefixes of existing keys
    and extensions of existing keys, each mode on top of the other."""
    rng = random.Random(71)
    keys = set()
    while len(keys) < 1200:
        keys.add(bytes(rng.choice([0, 1, 0x10, 0xfe]) for _ in range(rng.randint(1, 5))))
    keys = sorted(keys)
    rng.shuffle(keys)
    first, rest = keys[:700], keys[700:]
    second = rest + rng.sample(first, 200)
    rng.shuffle(second)
    t = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    for ks, mode in ((first, modes[0]), (second, modes[1])):
        vals = [bytes([rng.getrandbits(8)]) * rng.randint(1, 50) for _ in ks]
        b = workloads.batch_from_bytes(ks, vals)
        _put(t, b, mode)
        o.puts(b)
        o.hash()
        assert t.get_root_hash() == o.root()
    allb = workloads.batch_from_bytes(keys, [b"x"] * len(keys))
    assert t.get_values(allb) == o.gets(allb)


@pytest.mark.parametrize("mode", MODES)
def test_equal_length_key_packing_edges(gmpt, workloads, oracle, mode):
    """Batches whose keys all have key 0's length are packed by the validation
    pass (equal slots, sort windows read from the packed bytes); one key of
    another length anywhere, key 0 included, sends the batch down the general
    packing.  Cases: every key identical (the common prefix is the whole key),
    a shared prefix of 37 nibbles (the window starts mid-word), key 0 longer
    than the rest, one shorter key in the middle, and keys whose length is
    odd in nibbles."""
    rng = random.Random(71)

    def vals(n):
        return [bytes([rng.getrandbits(8)]) * rng.randint(1, 90) for _ in range(n)]

    same = [bytes(range(32))] * 500
    pre = bytes(rng.getrandbits(8) for _ in range(18)) + b"\x50"  # 37 shared nibbles + 1
    shared = list({pre + bytes(rng.getrandbits(8) for _ in range(13)) for _ in range(3000)})
    uni = list({bytes(rng.getrandbits(8) for _ in range(20)) for _ in range(3000)})
    long0 = [bytes(rng.getrandbits(8) for _ in range(21))] + uni
    mid = uni[:1500] + [uni[1500][:7]] + uni[1501:]
    cases = [("identical", same), ("shared-37", shared), ("key0-longer", long0), ("one-short", mid)]
    for name, keys in cases:
        b = workloads.batch_from_bytes(keys, vals(len(keys)))
        t = _build(gmpt, b, mode)
        assert t.get_root_hash() == oracle.root_of(b), name
        last = {}
        for i in range(b.n):
            last[b.key(i)] = b.value(i)
        assert t.get_values(b) == [last[b.key(i)] for i in range(b.n)], name
    # odd nibble counts: hex keys of 2 x 9 + 1 nibbles cut to 18 (17 + terminator)
    hk = list({bytes(rng.randrange(16) for _ in range(17)) + b"\x10" for _ in range(2500)})
    b = workloads.batch_from_hex(hk, vals(len(hk)))
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)


@pytest.mark.parametrize("mode", MODES)
def test_short_root_and_embedded_children(gmpt, workloads, oracle, mode):
    """The reference's encoding quirks on the GPU: a root whose encoding is
    under 32 bytes is returned raw (cpu_mpt.cuh:413-418, 915-926), and child
    encodings under 32 bytes are string-wrapped in their parent
    (node.cuh:75, 248, 254); the exported nodeset matches the oracle's."""
    b = workloads.batch_from_bytes([b"\x10", b"\x20"], [b"a", b"bb"])
    t = _build(gmpt, b, mode)
    root = t.get_root_hash()
    assert len(root) < 32 and root == oracle.root_of(b)
    o = oracle.OracleMPT()
    o.puts(b)
    o.hash()
    want = [(h + bytes(32 - len(h)), e) for h, e in o.export()]
    assert sorted(t.flush_dirty_nodes(b)) == sorted(want)
    assert t.get_values(b) == [b"a", b"bb"]
    # embedded children under a hashed root, then a second batch on top
    keys = [bytes([16 * k]) for k in range(1, 9)]
    b2 = workloads.batch_from_bytes(keys, [bytes([k]) * (k % 3) for k in range(8)])
    _put(t, b2, mode)
    o.puts(b2)
    o.hash()
    assert len(t.get_root_hash()) == 32 and t.get_root_hash() == o.root()


@pytest.mark.parametrize("mode", MODES)
def test_skewed_1000_byte_values_on_preloaded_trie(gmpt, workloads, oracle, mode):
    """The skewed-update shape with its full 1,000-byte values (leaves of
    eight Keccak blocks) at reduced size: a preloaded trie of 'user' +
    19-digit keys, then Zipf(0.99) overwrites (rewritten in place) and
    inserts, checked against the oracle after each batch."""
    pre, ops = workloads.config3_lock(seed=9, n_preload=60_000, n_ops=40_000, universe=200_000, value_len=1000)
    t = gmpt.MPT(device=0)
    _put(t, pre, "2phase")
    o = oracle.OracleMPT()
    o.puts(pre)
    o.hash()
    assert t.get_root_hash() == o.root()
    _put(t, ops, mode)
    o.puts(ops)
    o.hash()
    assert t.get_root_hash() == o.root()
    sample = ops.subset(range(0, ops.n, 53))
    assert t.get_values(sample) == o.gets(sample)


def _batched_shape(workloads, scale):
    """The batched workload's 4,096 lognormal trie sizes (median 256, sigma
    2.046, clipped to 1..200,000) divided by `scale` (at least one key each;
    the largest trie keeps its full size), 32-byte keys and values, the keys of
    all tries shuffled together."""
    sizes, rng = workloads.config4_sizes()
    big = int(np.argmax(sizes))
    sz = np.maximum(1, np.round(sizes / scale)).astype(np.int64)
    sz[big] = sizes[big]
    total = int(sz.sum())
    keys = workloads.unique_random_keys(rng, total, 32)
    vals = rng.integers(0, 256, size=(total, 32), dtype=np.uint8)
    tid = np.repeat(np.arange(len(sz), dtype=np.int32), sz)
    starts = np.concatenate([[0], np.cumsum(sz)[:-1]])
    perm = rng.permutation(total)
    allb = workloads.batch_from_fixed(keys[perm], vals[perm].reshape(-1), np.full(total, 32))
    allb.trie_ids = tid[perm]
    per = [workloads.batch_from_fixed(keys[s:s + c], vals[s:s + c].reshape(-1), np.full(int(c), 32))
           for s, c in zip(starts, sz)]
    return sz, big, allb, per


@pytest.mark.parametrize("mode", MODES)
def test_batched_4096_uneven_tries(gmpt, workloads, oracle, mode):
    """4,096 tries of uneven size (1-key tries up to one of 200,000 keys) in
    one multi-trie insert and one hash pass: every trie's root equals its
    plain single-trie build, and a sample including the largest trie equals
    the oracle's."""
    sz, big, allb, per = _batched_shape(workloads, 16)
    assert (sz == 1).sum() > 100 and sz[big] == 200_000
    t = gmpt.MPT(device=0, num_tries=len(sz))
    _put(t, allb, mode)
    roots = t.get_root_hashes()
    p = gmpt.MPT(device=0)
    for i, b in enumerate(per):
        p.clear()
        _put(p, b, "2phase")
        assert roots[i] == p.get_root_hash(), i
    rng = np.random.default_rng(4)
    for i in sorted(set(rng.choice(len(sz), 12, replace=False).tolist()) | {big, int(np.argmin(sz))}):
        assert roots[i] == oracle.root_of(per[i]), i
    t.close()
    p.close()


def test_batched_4096_uneven_tries_through_graph(gmpt, workloads, oracle):
    """The same 4,096-trie batch through the replayed-graph pass
    (gmpt_rebuild_graph): the capturing call and two replays give the eager
    multi-trie pass's roots, and a sample including the largest and a 1-key
    trie equals the oracle's."""
    sz, big, allb, per = _batched_shape(workloads, 16)
    e = gmpt.MPT(device=0, num_tries=len(sz))
    _put(e, allb, "2phase")
    want = e.get_root_hashes()
    e.close()
    db = gmpt.DeviceBatch(allb, adopt_values=True)
    t = gmpt.MPT(device=0, num_tries=len(sz))
    assert t.rebuild_graph(db) is False
    assert t.get_root_hashes() == want
    for _ in range(2):
        assert t.rebuild_graph(db) is True
        assert t.get_root_hashes() == want
    rng = np.random.default_rng(5)
    for i in sorted(set(rng.choice(len(sz), 8, replace=False).tolist()) | {big, int(np.argmin(sz))}):
        assert want[i] == oracle.root_of(per[i]), i
    t.close()
    db.free()


def test_graph_rebuild_replays_and_checks_its_batch(gmpt, workloads, oracle):
    """The batched pass as a replayed HIP graph (gmpt_rebuild_graph): the
    first call captures, later calls with the same buffers replay; roots equal
    the eager multi-trie pass's and the oracle's.  New contents of the same
    shape (other keys: other hash levels) fail the device's replay check and
    are redone eagerly with the right roots; the same contents again replay."""
    sz, big, allb, per = _batched_shape(workloads, 64)
    db = gmpt.DeviceBatch(allb, adopt_values=True)
    t = gmpt.MPT(device=0, num_tries=len(sz))
    e = gmpt.MPT(device=0, num_tries=len(sz))
    _put(e, allb, "2phase")
    want = e.get_root_hashes()
    assert t.rebuild_graph(db) is False  # captured (and replayed once)
    assert t.get_root_hashes() == want
    for _ in range(2):
        assert t.rebuild_graph(db) is True
        assert t.get_root_hashes() == want
    rng = np.random.default_rng(8)
    for i in sorted(set(rng.choice(len(sz), 6, replace=False).tolist()) | {big}):
        assert want[i] == oracle.root_of(per[i]), i
    # other keys and values in the same buffers (same lengths, index pairs and tries)
    keys2 = workloads.unique_random_keys(rng, allb.n, 32)
    vals2 = rng.integers(0, 256, size=(allb.n, 32), dtype=np.uint8)
    b2 = workloads.batch_from_fixed(keys2, vals2.reshape(-1), np.full(allb.n, 32))
    b2.trie_ids = allb.trie_ids
    assert np.array_equal(b2.keys_indexs, allb.keys_indexs) and np.array_equal(b2.values_indexs, allb.values_indexs)
    L = gmpt.gmpt.lib()
    for buf, arr in ((db.bufs[0], b2.keys_hexs), (db.bufs[2], b2.values_bytes)):
        assert L.gmpt_memcpy_htod(buf.ptr, arr.ctypes.data, arr.nbytes) == 0
    e.clear()
    _put(e, b2, "2phase")
    want2 = e.get_root_hashes()
    assert t.rebuild_graph(db) is False  # the replay check found other hash levels
    assert t.get_root_hashes() == want2
    assert t.rebuild_graph(db) is True
    assert t.get_root_hashes() == want2
    t.close()
    e.close()
    db.free()


def test_latching_overwrites_allocate_nothing(gmpt, workloads, oracle):
    """LockNU update-only batches rewrite the keys' value nodes in place and
    keep no packed keys: across five rounds of updates the pools grow by the
    staged value bytes only, and every root and value matches the oracle."""
    rng = np.random.default_rng(63)
    base = _rand_batch(workloads, 63, 4000, 30, 90)
    t = gmpt.MPT(device=0)
    o = oracle.OracleMPT()
    _put(t, base, "latching")
    o.puts(base)
    o.hash()
    assert t.get_root_hash() == o.root()
    used0 = t.stats()["pool_bytes_used"]
    staged = 0
    for r in range(5):
        flat, lens = workloads.random_values(rng, base.n, 20, 140)
        upd = workloads.batch_from_hex([base.key(i) for i in range(base.n)],
                                       [flat[int(lens[:i].sum()):int(lens[:i + 1].sum())].tobytes()
                                        for i in range(base.n)])
        upd = upd.subset(rng.permutation(upd.n))
        _put(t, upd, "latching")
        o.puts(upd)
        o.hash()
        assert t.get_root_hash() == o.root(), r
        staged += int(lens.sum()) + 8 + 256
    grown = t.stats()["pool_bytes_used"] - used0
    assert grown <= staged, (grown, staged)
    sample = base.subset(range(0, base.n, 13))
    assert t.get_values(sample) == o.gets(sample)
    # a batch of new keys and overwrites together still takes slots for the new ones only
    fresh = _rand_batch(workloads, 64, 3000, 30, 90)
    mix = workloads.concat_batches([fresh, base.subset(range(0, 2000))])
    _put(t, mix, "latching")
    o.puts(mix)
    o.hash()
    assert t.get_root_hash() == o.root()
    assert t.get_values(fresh) == o.gets(fresh)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("where", ["host", "device"])
def test_index_pairs_sharing_one_key(gmpt, workloads, oracle, mode, where):
    """10,000 index pairs that all point at one 65-nibble key (the packed keys
    need far more arena than the key buffer's size suggests): the arena is
    sized from the pairs (host batches) or grown after the packing pass
    reports the overflow (device batches), and the root is the oracle's."""
    key = bytes(range(32))
    hk = workloads.key_bytes_to_hex(key)
    n = 10_000
    rng = np.random.default_rng(5)
    flat, lens = workloads.random_values(rng, n, 30, 60)
    b = workloads.Batch(np.frombuffer(hk, np.uint8).copy(), np.tile(np.array([0, len(hk) - 1], np.int32), n),
                        flat.astype(np.uint8), workloads._index_pairs(lens, np.int64), n)
    t = gmpt.MPT(device=0)
    if where == "host":
        _put(t, b, mode)
    else:
        db = gmpt.DeviceBatch(b)
        fn = t.puts_2phase_with_valuehp if mode == "2phase" else t.puts_latching_with_valuehp_v2
        targets, nt = fn(db)
        t.hash_onepass_v2(targets, nt)
        db.free()
    assert t.get_root_hash() == oracle.root_of(b)
    assert t.get_values(b.subset([0])) == [b.value(n - 1)]


@pytest.mark.parametrize("mode", MODES)
def test_tail_beside_a_clean_long_leaf(gmpt, workloads, oracle, mode):
    """The deepest small levels are hashed on a second stream after their
    leaf children; a clean leaf child too long for the leaf word builder
    (a 20,000-byte value from an earlier batch, beside a key inserted now)
    keeps the hash its own pass gave it, and the root equals the oracle's."""
    rng = np.random.default_rng(81)
    def key(prefix):
        return bytes(prefix) + bytes(rng.integers(0, 256, 32 - len(prefix), dtype=np.uint8))
    k1, k2, k3 = key([0x00, 0x00]), key([0x00, 0x01]), key([0x00, 0x02])
    others = [key([int(b)]) for b in rng.integers(0x10, 0x100, 40)]
    big = bytes(rng.integers(0, 256, 20_000, dtype=np.uint8))
    first = workloads.batch_from_bytes([k1, k2] + others, [big, b"v2"] + [b"w%d" % i for i in range(len(others))])
    second = workloads.batch_from_bytes([k3], [b"v3"])
    t = gmpt.MPT(device=0)
    _put(t, first, mode)
    assert t.get_root_hash() == oracle.root_of(first)
    _put(t, second, mode)
    assert t.get_root_hash() == oracle.root_of(workloads.concat_batches([first, second]))
    t.close()


@pytest.mark.parametrize("mode", MODES)
def test_root_reads_between_other_calls(gmpt, workloads, oracle, mode):
    """A hash pass brings trie 0's root back with its closing counter read;
    every other call drops that copy, so a root read always shows the tries'
    current header: equal to the one-copy read of all roots, to the oracle's
    after each hash, unchanged by an insert that is not hashed yet, empty
    after a clear."""
    base = _rand_batch(workloads, 71, 3000, 1, 90)
    more = _rand_batch(workloads, 72, 1000, 1, 90)
    t = gmpt.MPT(device=0)
    _put(t, base, mode)
    r1 = t.get_root_hash()
    assert r1 == oracle.root_of(base)
    assert t.get_root_hash() == r1 and t.get_root_hashes(0, 1) == [r1]
    fn = t.puts_2phase_with_valuehp if mode == "2phase" else t.puts_latching_with_valuehp_v2
    targets, n = fn(more)
    assert t.get_root_hash() == r1 == t.get_root_hashes(0, 1)[0]  # the header changes with the hash only
    t.hash_onepass_v2(targets, n)
    r2 = t.get_root_hash()
    assert r2 == oracle.root_of(workloads.concat_batches([base, more])) and r2 != r1
    assert t.get_root_hashes(0, 1) == [r2]
    t.clear()
    assert t.get_root_hash() == b"" == t.get_root_hashes(0, 1)[0]
    _put(t, more, mode)
    assert t.get_root_hash() == oracle.root_of(more)
    t.close()


@pytest.mark.parametrize("mode", MODES)
def test_batch_of_more_than_4096_blocks(gmpt, workloads, oracle, mode):
    """1,100,000 keys in one batch: more than 4,096 blocks of 256, so the
    run numbering and the branch ranks take their block counts through a
    device scan instead of each block adding up the counts before it, and
    the split-point queries leave their LDS window for the widest
    intervals; the root equals the oracle's."""
    b = _rand_batch(workloads, 1101, 1_100_000, 1, 16)
    t = _build(gmpt, b, mode)
    assert t.get_root_hash() == oracle.root_of(b)
    t.close()
