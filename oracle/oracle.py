"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The C source
restates the reference's CPU MPT (mpt-with-compress/include/mpt/cpu_mpt.cuh);
see the header of oracle/mpt_oracle.c for the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_keccak256_pad.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint8, ctypes.c_void_p]
        L.orc_keccak_f1600.argtypes = [ctypes.c_void_p]
        L.orc_new.restype = ctypes.c_void_p
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_puts.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
        L.orc_puts.restype = ctypes.c_int
        L.orc_bulk_build.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
        L.orc_bulk_build.restype = ctypes.c_int
        L.orc_hash.argtypes = [ctypes.c_void_p]
        L.orc_root.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_root.restype = ctypes.c_int
        L.orc_gets.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_gets.restype = ctypes.c_int64
        L.orc_census.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_export.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
        L.orc_export.restype = ctypes.c_int64
        L.orc_hex_to_compact.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_hex_to_compact.restype = ctypes.c_int
        L.orc_rlp_write_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_rlp_write_bytes.restype = ctypes.c_int
        L.orc_rlp_puthead.argtypes = [ctypes.c_void_p, ctypes.c_uint8, ctypes.c_uint8, ctypes.c_uint64]
        L.orc_rlp_puthead.restype = ctypes.c_int
        L.orc_time_build.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_time_build.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def keccak256(data: bytes, pad: int = 0x01) -> bytes:
    buf = np.frombuffer(bytes(data), dtype=np.uint8) if data else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().orc_keccak256_pad(_ptr(buf), len(data), pad, _ptr(out))
    return out.tobytes()


def keccak_f1600(state):
    st = np.array(state, dtype=np.uint64)
    lib().orc_keccak_f1600(_ptr(st))
    return st


def hex_to_compact(hexs) -> bytes:
    h = np.array(list(hexs) or [0], dtype=np.uint8)
    out = np.zeros(len(hexs) // 2 + 2, np.uint8)
    n = lib().orc_hex_to_compact(_ptr(h), len(hexs), _ptr(out))
    return out[:n].tobytes()


def rlp_bytes(b: bytes) -> bytes:
    src = np.frombuffer(b, dtype=np.uint8) if b else np.zeros(1, np.uint8)
    out = np.zeros(len(b) + 16, np.uint8)
    n = lib().orc_rlp_write_bytes(_ptr(out), _ptr(src), len(b))
    return out[:n].tobytes()


def rlp_list_head(payload_len: int) -> bytes:
    out = np.zeros(16, np.uint8)
    n = lib().orc_rlp_puthead(_ptr(out), 0xC0, 0xF7, payload_len)
    return out[:n].tobytes()


class OracleMPT:
    """CpuMPT::Compress::MPT restated: puts_baseline / hashs_dirty_flag /
    get_root_hash / gets_baseline on one trie."""

    def __init__(self):
        self._h = lib().orc_new()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def puts(self, batch):
        rc = lib().orc_puts(self._h, _ptr(batch.keys_hexs), _ptr(batch.keys_indexs), _ptr(batch.values_bytes),
                            _ptr(batch.values_indexs), batch.n)
        if rc != 0:
            raise ValueError("oracle: malformed batch")

    def bulk_build(self, batch):
        rc = lib().orc_bulk_build(self._h, _ptr(batch.keys_hexs), _ptr(batch.keys_indexs), _ptr(batch.values_bytes),
                                  _ptr(batch.values_indexs), batch.n)
        if rc != 0:
            raise ValueError("oracle: duplicate keys in bulk build")

    def hash(self):
        lib().orc_hash(self._h)

    def root(self) -> bytes:
        out = np.zeros(32, np.uint8)
        n = lib().orc_root(self._h, _ptr(out))
        return out[:n].tobytes()

    def gets(self, batch):
        """Returns a list of bytes (or None for absent keys)."""
        n = batch.n
        sizes = np.zeros(max(n, 1), np.int32)
        offs = np.zeros(max(n, 1), np.int64)
        total = lib().orc_gets(self._h, _ptr(batch.keys_hexs), _ptr(batch.keys_indexs), n, _ptr(sizes), None, None)
        vals = np.zeros(max(total, 1), np.uint8)
        lib().orc_gets(self._h, _ptr(batch.keys_hexs), _ptr(batch.keys_indexs), n, _ptr(sizes), _ptr(vals), _ptr(offs))
        out = []
        for i in range(n):
            out.append(None if sizes[i] < 0 else vals[offs[i]:offs[i] + sizes[i]].tobytes())
        return out

    def census(self):
        out = np.zeros(8, np.int64)
        lib().orc_census(self._h, _ptr(out))
        keys = ["n_full", "n_short_leaf", "n_short_ext", "n_value_direct", "n_hashed", "n_embedded", "perms",
                "enc_bytes"]
        return dict(zip(keys, (int(v) for v in out)))

    def export(self):
        """All FULL/SHORT nodes as a list of (hash_or_raw_bytes, encoding_bytes)."""
        total = ctypes.c_int64(0)
        n = lib().orc_export(self._h, None, None, None, ctypes.byref(total))
        hashes = np.zeros(max(33 * n, 1), np.uint8)
        encs = np.zeros(max(total.value, 1), np.uint8)
        offs = np.zeros(max(2 * n, 1), np.int64)
        lib().orc_export(self._h, _ptr(hashes), _ptr(encs), _ptr(offs), ctypes.byref(total))
        out = []
        for i in range(n):
            hs = int(hashes[33 * i])
            out.append((hashes[33 * i + 1:33 * i + 1 + hs].tobytes(), encs[offs[2 * i]:offs[2 * i + 1] + 1].tobytes()))
        return out


def root_of(batch) -> bytes:
    t = OracleMPT()
    t.puts(batch)
    t.hash()
    return t.root()


def time_build(batch):
    """Seconds for puts + hash of batch into a fresh oracle trie, and the root."""
    out = np.zeros(32, np.uint8)
    sz = ctypes.c_int(0)
    secs = lib().orc_time_build(_ptr(batch.keys_hexs), _ptr(batch.keys_indexs), _ptr(batch.values_bytes),
                                _ptr(batch.values_indexs), batch.n, _ptr(out), ctypes.byref(sz))
    return secs, out[:sz.value].tobytes()
