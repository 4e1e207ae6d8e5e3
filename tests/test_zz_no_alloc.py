"""GPU checks of gmpt_build_root_noalloc, the entry point that builds a
trie's root in a fixed device region without allocating.

It is compared, through the C-ABI, with the existing entry points that build
a fresh trie's root -- init_mpt + insert_mpt_2phase / insert_mpt_olc
(mpt-with-compress/lib/libgmpt.cu:240-331) -- and with the oracle, on two
calls in a row with the second batch smaller, so a stale region left by the
first call would show."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand_batch(workloads, seed, n, vlo, vhi, key_len=32):
    rng = np.random.default_rng(seed)
    keys = workloads.unique_random_keys(rng, n, key_len)
    flat, lens = workloads.random_values(rng, n, vlo, vhi)
    return workloads.batch_from_fixed(keys, flat, lens)


def _fresh_trie_root(L, b, mode):
    """Root from init_mpt + insert_mpt_2phase (mode 0) / insert_mpt_olc (1)."""
    m = L.init_mpt()
    assert m
    args = [b.keys_hexs.ctypes.data, b.keys_indexs.ctypes.data, b.values_bytes.ctypes.data,
            b.values_indexs.ctypes.data, None, b.n]
    h = (L.insert_mpt_2phase if mode == 0 else L.insert_mpt_olc)(m, *args)
    assert h, L.gmpt_last_error().decode()
    root = ctypes.string_at(h, 32)
    L.gmpt_free_host(h)
    L.free_mpt(m)
    return root


@pytest.mark.parametrize("mode", [0, 1], ids=["2phase", "latching"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_noalloc_root_matches_fresh_trie(gmpt, workloads, oracle, mode, where):
    L = gmpt.gmpt.lib()
    big = _rand_batch(workloads, 501 + mode, 30000, 70, 110)
    small = _rand_batch(workloads, 601 + mode, 2500, 20, 300)
    for b in (big, small):
        src = gmpt.DeviceBatch(b) if where == "device" else b
        got = gmpt.build_root_noalloc(src, mode)
        assert got == _fresh_trie_root(L, b, mode)
        assert got == oracle.root_of(b)
        if where == "device":
            src.free()


@pytest.mark.parametrize("mode", [0, 1], ids=["2phase", "latching"])
def test_noalloc_duplicates_and_short_root(gmpt, workloads, oracle, mode):
    # last writer wins within the batch, as in the existing entry points
    rng = np.random.default_rng(77)
    keys = workloads.unique_random_keys(rng, 4000, 32)
    keys = np.concatenate([keys, keys[::3]])
    flat, lens = workloads.random_values(rng, len(keys), 33, 90)
    b = workloads.batch_from_fixed(keys, flat, lens)
    L = gmpt.gmpt.lib()
    got = gmpt.build_root_noalloc(b, mode)
    assert got == _fresh_trie_root(L, b, mode) == oracle.root_of(b)
    # a root encoding under 32 bytes comes back raw (cpu_mpt.cuh:915-926)
    s = workloads.batch_from_bytes([b"\x10", b"\x20"], [b"a", b"bb"])
    got = gmpt.build_root_noalloc(s, mode)
    assert len(got) < 32 and got == oracle.root_of(s)
    # and the empty batch has no root
    e = workloads.batch_from_bytes([], [])
    assert gmpt.build_root_noalloc(e, mode) == b""


def test_noalloc_rejects_batches_above_its_limit(gmpt, workloads):
    b = _rand_batch(workloads, 700, gmpt.gmpt.NOALLOC_MAX_KEYS + 1, 1, 2)
    with pytest.raises(gmpt.GmptError, match="limit"):
        gmpt.build_root_noalloc(b, 0)
    # values too large for the fixed region: a clean error, then a normal call
    big = _rand_batch(workloads, 701, 20000, 4000, 4100)
    with pytest.raises(gmpt.GmptError, match="scratch"):
        gmpt.build_root_noalloc(big, 0)
    ok = _rand_batch(workloads, 702, 100, 10, 20)
    assert len(gmpt.build_root_noalloc(ok, 1)) == 32
