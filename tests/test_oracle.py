"""Pins for the CPU oracle (oracle/mpt_oracle.c).

Each check is chosen so that a plausible slip in the oracle (a wrong pad byte,
rotation offset, RLP length rule, nibble order, trie split index, overwrite
rule) fails it: differential tests against hashlib's SHA3-256 (same
permutation and rate, pad 0x06), published Keccak/RLP/hex-prefix vectors, and
trie encodings derived by hand in this file from the RLP and hex-prefix rules
(independent Python code, not a call into the oracle).
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- Keccak ----

def test_keccak_f1600_zero_state_kat(oracle):
    # Keccak team's KeccakF-1600 intermediate values: permutation of the
    # all-zero state, first two lanes.
    st = oracle.keccak_f1600([0] * 25)
    assert int(st[0]) == 0xF1258F7940E1DDE7
    assert int(st[1]) == 0x84D5CCF933C0478A


def test_keccak_vs_sha3_differential(oracle):
    rng = random.Random(7)
    lengths = list(range(0, 300)) + [407, 408, 409, 543, 544, 545, 1000, 1087, 1088, 1089, 2000]
    for n in lengths:
        msg = bytes(rng.getrandbits(8) for _ in range(n))
        assert oracle.keccak256(msg, pad=0x06) == hashlib.sha3_256(msg).digest(), n


def test_keccak256_kats(oracle):
    # Keccak-256 of the empty string, and of RLP("") = 0x80, Ethereum's empty-trie root.
    assert oracle.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"
    assert oracle.keccak256(b"\x80").hex() == "56e81f171bcc55a6ff8345e692c0f86e5b48e01b996cadc001622fb5e363b421"
    assert oracle.keccak256(b"abc").hex() == "4e03657aea45a94fc7d47ba826c8d667c0d1e6e33a64a036ec44f58fa12d6c45"


def test_keccak_golden_vectors(oracle):
    with open(os.path.join(GOLDEN, "keccak256.json")) as f:
        vecs = json.load(f)
    for v in vecs:
        assert oracle.keccak256(bytes.fromhex(v["msg"])).hex() == v["digest"]


# ------------------------------------------------------------------- RLP ----

def test_rlp_vectors(oracle):
    assert oracle.rlp_bytes(b"dog") == bytes.fromhex("83646f67")
    assert oracle.rlp_bytes(b"") == b"\x80"
    assert oracle.rlp_bytes(b"\x0f") == b"\x0f"
    assert oracle.rlp_bytes(b"\x7f") == b"\x7f"
    assert oracle.rlp_bytes(b"\x80") == b"\x81\x80"
    s56 = b"Lorem ipsum dolor sit amet, consectetur adipisicing elit"
    assert len(s56) == 56
    assert oracle.rlp_bytes(s56) == b"\xb8\x38" + s56
    s1024 = bytes(1024)
    assert oracle.rlp_bytes(s1024)[:3] == b"\xb9\x04\x00"
    assert oracle.rlp_list_head(0) == b"\xc0"
    assert oracle.rlp_list_head(8) + oracle.rlp_bytes(b"cat") + oracle.rlp_bytes(b"dog") == \
        bytes.fromhex("c88363617483646f67")
    assert oracle.rlp_list_head(55) == b"\xf7"
    assert oracle.rlp_list_head(56) == b"\xf8\x38"
    assert oracle.rlp_list_head(532) == b"\xf9\x02\x14"


def test_hex_prefix_vectors(oracle):
    # go-ethereum trie/encoding_test.go vectors, also held (commented out) at
    # mpt-with-compress/test/utils.cu:42-56.
    cases = [([], "00"), ([16], "20"), ([1, 2, 3, 4, 5], "112345"), ([0, 1, 2, 3, 4, 5], "00012345"),
             ([15, 1, 12, 11, 8, 16], "3f1cb8"), ([0, 15, 1, 12, 11, 8, 16], "200f1cb8")]
    for hexs, want in cases:
        assert oracle.hex_to_compact(hexs).hex() == want, hexs


# ---------------------------------------------- hand-derived trie encodings --

def _rlp_str(b):
    if len(b) == 1 and b[0] < 0x80:
        return b
    if len(b) < 56:
        return bytes([0x80 + len(b)]) + b
    ln = len(b).to_bytes((len(b).bit_length() + 7) // 8, "big")
    return bytes([0xB7 + len(ln)]) + ln + b


def _rlp_list(payload):
    if len(payload) < 56:
        return bytes([0xC0 + len(payload)]) + payload
    ln = len(payload).to_bytes((len(payload).bit_length() + 7) // 8, "big")
    return bytes([0xF7 + len(ln)]) + ln + payload


def _hp(nibbles, leaf):
    flag = 2 if leaf else 0
    if len(nibbles) % 2:
        out = [(flag + 1) << 4 | nibbles[0]]
        nibbles = nibbles[1:]
    else:
        out = [flag << 4]
    for i in range(0, len(nibbles), 2):
        out.append(nibbles[i] << 4 | nibbles[i + 1])
    return bytes(out)


def _nib(key):
    out = []
    for b in key:
        out += [b >> 4, b & 15]
    return out


def _ref(enc, oracle):
    """What a parent stores for a child: digest if enc >= 32 B, else the raw encoding."""
    return oracle.keccak256(enc) if len(enc) >= 32 else enc


def test_single_key_root(oracle, workloads):
    k = bytes(range(32))
    v = bytes(range(100, 164))
    b = workloads.batch_from_bytes([k], [v])
    enc = _rlp_list(_rlp_str(_hp(_nib(k), True)) + _rlp_str(v))
    assert oracle.root_of(b) == oracle.keccak256(enc)


def test_two_keys_branch_root(oracle, workloads):
    k1 = bytes([0x12]) + bytes(31)
    k2 = bytes([0xA0]) + bytes(range(31))
    v1, v2 = b"\x01" * 40, b"\x02" * 50
    b = workloads.batch_from_bytes([k2, k1], [v2, v1])
    l1 = _rlp_list(_rlp_str(_hp(_nib(k1)[1:], True)) + _rlp_str(v1))
    l2 = _rlp_list(_rlp_str(_hp(_nib(k2)[1:], True)) + _rlp_str(v2))
    items = [b"\x80"] * 17
    items[0x1] = _rlp_str(_ref(l1, oracle))
    items[0xA] = _rlp_str(_ref(l2, oracle))
    root = _rlp_list(b"".join(items))
    assert oracle.root_of(b) == oracle.keccak256(root)


def test_extension_root(oracle, workloads):
    # three keys sharing the nibbles 'abc', diverging at nibble 3, two of them
    # sharing one more nibble -> ext(abc) -> branch{1: leaf, 7: ext? no: branch}
    k1 = bytes([0xAB, 0xC1]) + bytes(30)
    k2 = bytes([0xAB, 0xC7, 0x10]) + bytes(29)
    k3 = bytes([0xAB, 0xC7, 0x20]) + bytes(29)
    vs = [bytes([i]) * (40 + i) for i in range(3)]
    b = workloads.batch_from_bytes([k1, k2, k3], vs)
    n1, n2, n3 = _nib(k1), _nib(k2), _nib(k3)
    leaf1 = _rlp_list(_rlp_str(_hp(n1[4:], True)) + _rlp_str(vs[0]))
    # k2, k3 share 'abc7' then branch at depth 4 on nibbles 1 / 2
    leaf2 = _rlp_list(_rlp_str(_hp(n2[5:], True)) + _rlp_str(vs[1]))
    leaf3 = _rlp_list(_rlp_str(_hp(n3[5:], True)) + _rlp_str(vs[2]))
    inner = [b"\x80"] * 17
    inner[1] = _rlp_str(_ref(leaf2, oracle))
    inner[2] = _rlp_str(_ref(leaf3, oracle))
    inner_enc = _rlp_list(b"".join(inner))
    br = [b"\x80"] * 17
    br[1] = _rlp_str(_ref(leaf1, oracle))
    br[7] = _rlp_str(_ref(inner_enc, oracle))
    br_enc = _rlp_list(b"".join(br))
    ext = _rlp_list(_rlp_str(_hp([0xA, 0xB, 0xC], False)) + _rlp_str(_ref(br_enc, oracle)))
    assert oracle.root_of(b) == oracle.keccak256(ext)


def test_embedded_child_is_string_wrapped(oracle, workloads):
    # Reference quirk (node.cuh:75,248,254): a child whose encoding is < 32 B is
    # written with write_bytes (string-wrapped), not inlined as a list.
    k1, k2 = b"\x10", b"\x20"
    v1, v2 = b"a", b"bb"
    b = workloads.batch_from_bytes([k1, k2], [v1, v2])
    l1 = _rlp_list(_rlp_str(_hp([0], True)) + _rlp_str(v1))
    l2 = _rlp_list(_rlp_str(_hp([0], True)) + _rlp_str(v2))
    assert len(l1) < 32 and len(l2) < 32
    items = [b"\x80"] * 17
    items[1] = _rlp_str(l1)
    items[2] = _rlp_str(l2)
    root = _rlp_list(b"".join(items))
    # root encoding < 32 B: the reference returns it raw (cpu_mpt.cuh:413-418, 915-926)
    assert len(root) < 32
    assert oracle.root_of(b) == root


# ------------------------------------------------------------ invariants ----

def test_reference_doe_dog_gets(oracle, workloads):
    # mpt-with-compress/test/trie.cu:153-197 (CpuMpt.PutsBaselineBasic)
    b = workloads.batch_from_bytes([b"doe", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    t = oracle.OracleMPT()
    t.puts(b)
    assert t.gets(b) == [b"reindeer", b"puppy", b"cat"]
    miss = workloads.batch_from_bytes([b"do", b"dogg", b"cat"], [b"", b"", b""])
    assert t.gets(miss) == [None, None, None]


def test_reference_override_last_writer_wins(oracle, workloads):
    # mpt-with-compress/test/trie.cu:293-328 (CpuMpt.PutsBaselineOverride)
    b = workloads.batch_from_bytes([b"dog", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    t = oracle.OracleMPT()
    t.puts(b)
    assert t.gets(b) == [b"puppy", b"puppy", b"cat"]


def test_order_independence_and_bulk(oracle, workloads):
    rng = np.random.default_rng(11)
    keys = workloads.unique_random_keys(rng, 3000, 32)
    flat, lens = workloads.random_values(rng, 3000, 33, 150)
    b = workloads.batch_from_fixed(keys, flat, lens)
    r0 = oracle.root_of(b)
    perm = rng.permutation(b.n)
    assert oracle.root_of(b.subset(perm)) == r0
    t = oracle.OracleMPT()
    t.bulk_build(b)
    t.hash()
    assert t.root() == r0


def test_variable_length_keys_bulk_vs_incremental(oracle, workloads):
    rng = random.Random(5)
    keys = set()
    while len(keys) < 400:
        keys.add(bytes(rng.getrandbits(8) % 4 for _ in range(rng.randint(1, 5))))
    keys = sorted(keys)
    vals = [bytes([rng.getrandbits(8)]) * rng.randint(1, 60) for _ in keys]
    b = workloads.batch_from_bytes(keys, vals)
    t1 = oracle.OracleMPT()
    t1.puts(b.subset(np.random.default_rng(1).permutation(b.n)))
    t1.hash()
    t2 = oracle.OracleMPT()
    t2.bulk_build(b)
    t2.hash()
    assert t1.root() == t2.root()
    assert sorted(t1.export()) == sorted(t2.export())


def test_incremental_batches_equal_one_batch(oracle, workloads):
    rng = np.random.default_rng(3)
    keys = workloads.unique_random_keys(rng, 2000, 32)
    flat, lens = workloads.random_values(rng, 2000, 33, 120)
    b = workloads.batch_from_fixed(keys, flat, lens)
    t = oracle.OracleMPT()
    t.puts(b.subset(np.arange(0, 700)))
    t.hash()
    t.puts(b.subset(np.arange(700, 2000)))
    t.hash()
    assert t.root() == oracle.root_of(b)


def test_duplicates_equal_last_occurrence(oracle, workloads):
    rng = np.random.default_rng(9)
    keys = workloads.unique_random_keys(rng, 500, 32)
    idx = rng.integers(0, 500, size=1500)
    flat, lens = workloads.random_values(rng, 1500, 40, 90)
    b = workloads.batch_from_fixed(keys[idx], flat, lens)
    last = {}
    for i in range(b.n):
        last[b.key(i)] = i
    dedup = b.subset(sorted(last.values()))
    assert oracle.root_of(b) == oracle.root_of(dedup)


def test_same_value_update_keeps_root(oracle, workloads):
    rng = np.random.default_rng(4)
    keys = workloads.unique_random_keys(rng, 300, 32)
    flat, lens = workloads.random_values(rng, 300, 40, 90)
    b = workloads.batch_from_fixed(keys, flat, lens)
    t = oracle.OracleMPT()
    t.puts(b)
    t.hash()
    r = t.root()
    t.puts(b.subset(np.arange(0, 50)))
    t.hash()
    assert t.root() == r


def test_config1_golden_root(oracle, workloads):
    with open(os.path.join(GOLDEN, "roots.json")) as f:
        g = json.load(f)
    b = workloads.config1_oracle()
    assert oracle.root_of(b).hex() == g["config1"]


def test_census_config1(oracle, workloads):
    b = workloads.config1_oracle()
    t = oracle.OracleMPT()
    t.puts(b)
    t.hash()
    c = t.census()
    assert c["n_short_leaf"] == b.n
    assert c["n_hashed"] == c["n_full"] + c["n_short_leaf"] + c["n_short_ext"]
    assert c["n_embedded"] == 0


# ------------------------------------------------------------ node export ----

def _rlp_decode(b, i=0):
    """One RLP item at b[i:] -> (bytes or list, next index); strict lengths."""
    p = b[i]
    if p < 0x80:
        return bytes([p]), i + 1
    if p < 0xC0:
        if p <= 0xB7:
            n, s = p - 0x80, i + 1
        else:
            ll = p - 0xB7
            n, s = int.from_bytes(b[i + 1:i + 1 + ll], "big"), i + 1 + ll
        assert s + n <= len(b)
        return bytes(b[s:s + n]), s + n
    if p <= 0xF7:
        n, s = p - 0xC0, i + 1
    else:
        ll = p - 0xF7
        n, s = int.from_bytes(b[i + 1:i + 1 + ll], "big"), i + 1 + ll
    end, items = s + n, []
    assert end <= len(b)
    while s < end:
        it, s = _rlp_decode(b, s)
        items.append(it)
    assert s == end
    return items, end


def _rlp_encode(x):
    if isinstance(x, list):
        return _rlp_list(b"".join(_rlp_encode(y) for y in x))
    return _rlp_str(x)


def _hp_decode(c):
    """Hex-prefix (compact) bytes -> (nibbles, leaf flag)."""
    flag = c[0] >> 4
    assert flag < 4
    nib = [c[0] & 15] if flag & 1 else []
    if not flag & 1:
        assert c[0] & 15 == 0
    for x in c[1:]:
        nib += [x >> 4, x & 15]
    return nib, bool(flag & 2)


def _nodeset_contents(nodes, root, oracle):
    """Walks an exported nodeset from the root: every (h, e) must satisfy
    h == keccak256(e) (e >= 32 B) or h == e, every e must be an RLP list of 2
    or 17 items that re-encodes to itself, and the walk must visit each
    exported entry exactly once (identical subtrees, such as equal embedded
    leaves under different branches, are exported once per occurrence).
    Returns {key bytes: value bytes} read off the leaves."""
    by_hash, listed = {}, {}
    for h, e in nodes:
        assert h == (oracle.keccak256(e) if len(e) >= 32 else e)
        item, end = _rlp_decode(e)
        assert end == len(e) and isinstance(item, list) and len(item) in (2, 17)
        assert _rlp_encode(item) == e
        by_hash[h] = item
        listed[h] = listed.get(h, 0) + 1
    assert root in by_hash
    raw, visits = {}, {}

    def visit(ref, path):
        assert ref in by_hash
        visits[ref] = visits.get(ref, 0) + 1
        item = by_hash[ref]
        if len(item) == 17:
            for s in range(16):
                if item[s] != b"":
                    visit(item[s], path + (s,))
            if item[16] != b"":
                raw[path] = item[16]
            return
        nib, leaf = _hp_decode(item[0])
        if leaf:
            raw[path + tuple(nib)] = item[1]
        else:
            assert nib
            visit(item[1], path + tuple(nib))

    visit(root, ())
    assert visits == listed  # every exported node is in the trie, once per occurrence
    out = {}
    for p, v in raw.items():
        assert len(p) % 2 == 0 and p not in out
        out[bytes(p[k] << 4 | p[k + 1] for k in range(0, len(p), 2))] = v
    return out


def _export_cases(workloads):
    rng = np.random.default_rng(21)
    keys = workloads.unique_random_keys(rng, 2500, 32)
    flat, lens = workloads.random_values(rng, 2500, 1, 140)
    yield "random", workloads.batch_from_fixed(keys, flat, lens)
    r = random.Random(22)
    ks = set()
    while len(ks) < 600:  # prefixes of one another: values in branch slot 16
        ks.add(bytes(r.choice([0, 1, 0x10, 0xfe]) for _ in range(r.randint(1, 5))))
    ks = sorted(ks)
    r.shuffle(ks)
    yield "prefix-keys", workloads.batch_from_bytes(ks, [bytes([r.getrandbits(8)]) * r.randint(1, 50) for _ in ks])
    ks = list({bytes(r.getrandbits(8) for _ in range(r.randint(1, 2))) for _ in range(300)})
    yield "embedded", workloads.batch_from_bytes(ks, [bytes(r.getrandbits(8) for _ in range(r.randint(0, 3)))
                                                      for _ in ks])
    yield "short-root", workloads.batch_from_bytes([b"\x10", b"\x20"], [b"a", b"bb"])


def test_export_nodeset_is_the_trie(oracle, workloads):
    """orc_export (the oracle's flush_dirty_nodes / get_all_nodes nodeset):
    decoding the exported nodes from the root gives back exactly the batch's
    last-written key -> value map, each hash slot is Keccak-256 of its
    encoding (or the encoding itself under 32 bytes), every encoding is a
    well-formed RLP list of 2 or 17 items, and the node count equals the
    census' branch + short count."""
    for name, b in _export_cases(workloads):
        t = oracle.OracleMPT()
        t.puts(b)
        t.hash()
        c = t.census()
        nodes = t.export()
        assert len(nodes) == c["n_full"] + c["n_short_leaf"] + c["n_short_ext"], name
        want = {}
        for i in range(b.n):
            hk = b.key(i)
            assert hk[-1] == 16 and len(hk) % 2 == 1
            want[bytes(hk[k] << 4 | hk[k + 1] for k in range(0, len(hk) - 1, 2))] = b.value(i)
        # an empty value in a branch's slot 16 (a key that is a prefix of
        # another) encodes as 0x80, the same as an empty slot
        want = {k: v for k, v in want.items() if v or not any(o != k and o.startswith(k) for o in want)}
        assert _nodeset_contents(nodes, t.root(), oracle) == want, name
