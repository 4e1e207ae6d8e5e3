"""Batch layout and seeded synthetic workloads.

A batch uses the reference's packed layout (mpt-with-compress/include/util/
utils.cuh:254-309): ``keys_hexs`` holds every key as hex nibbles (one byte per
nibble, terminator 16 last, as key_bytes_to_hex writes them at utils.cuh:340-349),
``keys_indexs`` holds an inclusive ``[start, end]`` int32 pair per key, and
``values_bytes`` / ``values_indexs`` the same for values with int64 pairs.

The generators below produce the five benchmark workloads (sizes and
distributions as stated in each docstring) from numpy's PCG64 with a fixed
seed, so every run of a test or of the benchmark sees the same bytes.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class Batch:
    keys_hexs: np.ndarray      # uint8 nibbles
    keys_indexs: np.ndarray    # int32, 2 per key, inclusive
    values_bytes: np.ndarray   # uint8
    values_indexs: np.ndarray  # int64, 2 per value, inclusive
    n: int
    trie_ids: Optional[np.ndarray] = None  # int32 per key (multi-trie batches)

    def key(self, i) -> bytes:
        return self.keys_hexs[self.keys_indexs[2 * i]:self.keys_indexs[2 * i + 1] + 1].tobytes()

    def value(self, i) -> bytes:
        return self.values_bytes[self.values_indexs[2 * i]:self.values_indexs[2 * i + 1] + 1].tobytes()

    def subset(self, idx):
        """A new batch holding keys idx (in that order)."""
        idx = np.asarray(idx, dtype=np.int64)
        keys = [self.key(int(i)) for i in idx]
        vals = [self.value(int(i)) for i in idx]
        b = batch_from_hex(keys, vals)
        if self.trie_ids is not None:
            b.trie_ids = self.trie_ids[idx].astype(np.int32)
        return b

    @property
    def key_bytes(self):
        return int(self.keys_indexs[-1]) + 1 if self.n else 0

    @property
    def value_bytes(self):
        return int(self.values_indexs[-1]) + 1 if self.n else 0


def _concat(parts, dtype):
    if not parts:
        return np.zeros(0, dtype)
    return np.concatenate([np.frombuffer(p, dtype=np.uint8) if isinstance(p, (bytes, bytearray)) else
                           np.asarray(p, dtype=np.uint8) for p in parts]).astype(dtype, copy=False)


def _index_pairs(lengths, dtype):
    lengths = np.asarray(lengths, dtype=np.int64)
    ends = np.cumsum(lengths) - 1
    starts = ends - lengths + 1
    out = np.empty(2 * len(lengths), dtype=dtype)
    out[0::2] = starts
    out[1::2] = ends
    return out


def key_bytes_to_hex(key: bytes) -> bytes:
    """utils.cuh:340-349: high nibble first, terminator 16 appended."""
    out = bytearray()
    for b in key:
        out.append(b >> 4)
        out.append(b & 15)
    out.append(16)
    return bytes(out)


def batch_from_hex(keys_hex, values) -> Batch:
    keys_hex = [bytes(k) for k in keys_hex]
    values = [bytes(v) for v in values]
    kh = _concat(keys_hex, np.uint8)
    vb = _concat(values, np.uint8)
    return Batch(kh, _index_pairs([len(k) for k in keys_hex], np.int32), vb,
                 _index_pairs([len(v) for v in values], np.int64), len(keys_hex))


def batch_from_bytes(keys, values) -> Batch:
    return batch_from_hex([key_bytes_to_hex(k) for k in keys], values)


def batch_from_fixed(keys: np.ndarray, values_flat: np.ndarray, value_lengths: np.ndarray) -> Batch:
    """Vectorised: keys is (n, L) uint8 of raw key bytes."""
    n, L = keys.shape
    hexs = np.empty((n, 2 * L + 1), np.uint8)
    hexs[:, 0:2 * L:2] = keys >> 4
    hexs[:, 1:2 * L:2] = keys & 15
    hexs[:, 2 * L] = 16
    return Batch(hexs.reshape(-1), _index_pairs(np.full(n, 2 * L + 1), np.int32), values_flat.astype(np.uint8),
                 _index_pairs(value_lengths, np.int64), n)


def concat_batches(batches, trie_ids=None) -> Batch:
    """Concatenate batches; trie_ids (one per batch) tags the keys of each."""
    keys = [b.keys_hexs for b in batches]
    vals = [b.values_bytes for b in batches]
    klen = np.concatenate([np.diff(b.keys_indexs.reshape(-1, 2), axis=1)[:, 0] + 1 for b in batches])
    vlen = np.concatenate([np.diff(b.values_indexs.reshape(-1, 2), axis=1)[:, 0] + 1 for b in batches])
    out = Batch(np.concatenate(keys).astype(np.uint8), _index_pairs(klen, np.int32),
                np.concatenate(vals).astype(np.uint8), _index_pairs(vlen, np.int64), int(len(klen)))
    if trie_ids is not None:
        out.trie_ids = np.concatenate([np.full(b.n, t, np.int32) for b, t in zip(batches, trie_ids)])
    return out


def unique_random_keys(rng, n, key_len=32) -> np.ndarray:
    """n distinct uniformly random keys of key_len bytes (rejection of repeats)."""
    keys = rng.integers(0, 256, size=(n, key_len), dtype=np.uint8)
    while True:
        v = np.ascontiguousarray(keys).view(np.dtype((np.void, key_len))).reshape(-1)
        _, first = np.unique(v, return_index=True)
        if len(first) == n:
            return keys
        dup = np.setdiff1d(np.arange(n), first)
        keys[dup] = rng.integers(0, 256, size=(len(dup), key_len), dtype=np.uint8)


def random_values(rng, n, lo, hi):
    """Values of uniform random length in [lo, hi] (inclusive) of random bytes."""
    lens = rng.integers(lo, hi + 1, size=n).astype(np.int64)
    flat = rng.integers(0, 256, size=int(lens.sum()), dtype=np.uint8)
    return flat, lens


def config1_oracle(seed=1, n=10_000) -> Batch:
    """10,000 unique uniform 32-byte keys, 64-byte random values."""
    rng = np.random.Generator(np.random.PCG64(seed))
    keys = unique_random_keys(rng, n, 32)
    flat, lens = random_values(rng, n, 64, 64)
    return batch_from_fixed(keys, flat, lens)


def config2_phase(seed=2, n=1_000_000) -> Batch:
    """1M unique uniform 32-byte keys, values of uniform length 70..110 bytes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    keys = unique_random_keys(rng, n, 32)
    flat, lens = random_values(rng, n, 70, 110)
    return batch_from_fixed(keys, flat, lens)


# --- Zipf (Gray et al. "Quickly generating billion-record synthetic databases",
#     the generator YCSB's ZipfianGenerator uses) -------------------------------

def _zeta(n, theta):
    i = np.arange(1, n + 1, dtype=np.float64)
    return float(np.sum(i ** -theta))


def zipf_ranks(rng, n_items, theta, count):
    zetan = _zeta(n_items, theta)
    zeta2 = 1.0 + 0.5 ** theta
    alpha = 1.0 / (1.0 - theta)
    eta = (1.0 - (2.0 / n_items) ** (1.0 - theta)) / (1.0 - zeta2 / zetan)
    u = rng.random(count)
    uz = u * zetan
    r = (n_items * (eta * u - eta + 1.0) ** alpha).astype(np.int64)
    r = np.where(uz < 1.0 + 0.5 ** theta, 1, r)
    r = np.where(uz < 1.0, 0, r)
    return np.clip(r, 0, n_items - 1)


def user_keys(ids) -> np.ndarray:
    """'user' + 19-digit zero-padded decimal id: 23 ASCII bytes per key."""
    ids = np.asarray(ids, dtype=np.int64)
    out = np.empty((len(ids), 23), np.uint8)
    out[:, :4] = np.frombuffer(b"user", np.uint8)
    v = ids.copy()
    for d in range(22, 3, -1):
        out[:, d] = 48 + (v % 10)
        v //= 10
    return out


def config3_lock(seed=3, n_preload=4_000_000, n_ops=1_000_000, universe=10_000_000, value_len=1000,
                 theta=0.99):
    """Preload n_preload keys, then n_ops ops: a fair coin picks overwrite
    (Zipf rank over the preloaded set) or insert (Zipf rank over the rest of
    the id universe).  Keys 'user%019d', 1,000-byte random values.
    Returns (preload_batch, ops_batch)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(universe)
    pre, fresh = perm[:n_preload], perm[n_preload:]
    pre_keys = user_keys(pre)
    pre_vals = rng.integers(0, 256, size=n_preload * value_len, dtype=np.uint8)
    preload = batch_from_fixed(pre_keys, pre_vals, np.full(n_preload, value_len))
    coin = rng.random(n_ops) < 0.5
    r_over = zipf_ranks(rng, len(pre), theta, n_ops)
    r_ins = zipf_ranks(rng, len(fresh), theta, n_ops)
    ids = np.where(coin, pre[r_over], fresh[r_ins])
    op_vals = rng.integers(0, 256, size=n_ops * value_len, dtype=np.uint8)
    ops = batch_from_fixed(user_keys(ids), op_vals, np.full(n_ops, value_len))
    return preload, ops


def config4_sizes(seed=4, tries=4096, median=256, sigma=2.046, cap=200_000):
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.standard_normal(tries)
    return np.clip(np.round(np.exp(np.log(median) + sigma * z)), 1, cap).astype(np.int64), rng


def config4_batched(seed=4, tries=4096, median=256, sigma=2.046, cap=200_000, value_len=32):
    """`tries` independent tries, entry counts lognormal(median, sigma) clipped
    to [1, cap]; 32-byte uniform keys, 32-byte values.  One multi-trie batch
    with trie_ids."""
    sizes, rng = config4_sizes(seed, tries, median, sigma, cap)
    total = int(sizes.sum())
    keys = unique_random_keys(rng, total, 32)
    vals = rng.integers(0, 256, size=total * value_len, dtype=np.uint8)
    b = batch_from_fixed(keys, vals, np.full(total, value_len))
    b.trie_ids = np.repeat(np.arange(tries, dtype=np.int32), sizes)
    return b, sizes


def served_requests(seed=5, n_requests=1000, tries=1024, theta=0.8, median=2000, sigma=1.0, cap=50_000,
                    vlo=70, vhi=110):
    """Request stream for the serving workload: each request targets one trie
    (Zipf(theta) over `tries`) with a batch of lognormal(median, sigma) size
    clipped to [1, cap] of 32-byte uniform keys and 70..110-byte values."""
    rng = np.random.Generator(np.random.PCG64(seed))
    targets = zipf_ranks(rng, tries, theta, n_requests)
    sizes = np.clip(np.round(np.exp(np.log(median) + sigma * rng.standard_normal(n_requests))), 1, cap)
    sizes = sizes.astype(np.int64)
    reqs = []
    for t, s in zip(targets, sizes):
        keys = rng.integers(0, 256, size=(int(s), 32), dtype=np.uint8)
        flat, lens = random_values(rng, int(s), vlo, vhi)
        reqs.append((int(t), batch_from_fixed(keys, flat, lens)))
    return reqs
