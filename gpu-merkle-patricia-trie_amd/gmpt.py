"""ctypes binding of libgmpt_amd.so (include/gmpt_amd.h, include/libgmpt.h).

``MPT`` mirrors the reference's host class GpuMPT::Compress::MPT
(mpt-with-compress/include/mpt/gpu_mpt.cuh:18-192): the same method names and
argument meaning, with batches passed as ``workloads.Batch``.  Every call goes
through the C-ABI into the HIP kernels; there is no CPU path, and loading
fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgmpt_amd.so")
_lib = None


class GmptError(RuntimeError):
    pass


class _Batch(ctypes.Structure):
    _fields_ = [("keys_hexs", ctypes.c_void_p), ("keys_indexs", ctypes.c_void_p),
                ("values_bytes", ctypes.c_void_p), ("values_indexs", ctypes.c_void_p),
                ("values_hps", ctypes.c_void_p), ("trie_ids", ctypes.c_void_p), ("n", ctypes.c_int),
                ("keys_bytes", ctypes.c_int64), ("values_bytes_total", ctypes.c_int64),
                ("on_device", ctypes.c_int), ("adopt_values", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("last_hash_ms", ctypes.c_double), ("hash_launches", ctypes.c_uint32),
                ("hashed_nodes", ctypes.c_uint32), ("restarts", ctypes.c_uint32), ("lock_fails", ctypes.c_uint32),
                ("fallback_sorts", ctypes.c_uint32), ("regroups", ctypes.c_uint32), ("perms", ctypes.c_uint64),
                ("enc_bytes", ctypes.c_uint64), ("pool_bytes_used", ctypes.c_uint64),
                ("pool_bytes_reserved", ctypes.c_uint64), ("leaf_perm_ms", ctypes.c_double),
                ("leaf_perms", ctypes.c_uint64), ("leaf_bytes", ctypes.c_uint64), ("leaf_nodes", ctypes.c_uint32),
                ("ext_direct_nodes", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol the headers in include/ declare
GMPT_SYMBOLS = ["gmpt_forest_create", "gmpt_forest_destroy", "gmpt_last_error", "gmpt_puts_2phase",
                "gmpt_puts_latching", "gmpt_puts_with_read", "gmpt_hash_onepass", "gmpt_get_root_hash", "gmpt_get_root_hashes", "gmpt_gets", "gmpt_get_values",
                "gmpt_export_nodes", "gmpt_free_host", "gmpt_reset_trie", "gmpt_forest_clear", "gmpt_synchronize",
                "gmpt_set_option", "gmpt_get_option", "gmpt_autotune", "gmpt_get_stats",
                "gmpt_device_count", "gmpt_device_malloc", "gmpt_device_free", "gmpt_memcpy_htod",
                "gmpt_memcpy_dtoh", "gmpt_device_copy", "gmpt_timer_start", "gmpt_timer_stop", "gmpt_keccak256_batch",
                "gmpt_keccak256_batch_mode", "gmpt_build_root_noalloc", "gmpt_rebuild_graph", "gmpt_server_create", "gmpt_server_destroy", "gmpt_server_submit",
                "gmpt_server_poll", "gmpt_server_wait", "gmpt_server_drain", "gmpt_server_get_stats",
                "gmpt_server_latencies"]
LIBGMPT_SYMBOLS = ["build_mpt_2phase", "build_mpt_olc", "preprocess", "get_all_nodes", "init_mpt",
                   "insert_mpt_2phase", "insert_mpt_olc", "free_mpt"]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("batched_requests", ctypes.c_uint64),
                ("batched_keys", ctypes.c_uint64), ("deferred_mem", ctypes.c_uint64), ("busy_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _NodeSet(ctypes.Structure):
    _fields_ = [("hashs", ctypes.c_void_p), ("encs", ctypes.c_void_p), ("encs_indexs", ctypes.c_void_p),
                ("num", ctypes.c_ulonglong)]


def lib():
    """Load libgmpt_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GmptError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, u32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int64
    sig = {
        "gmpt_forest_create": ([i32, u32, ctypes.POINTER(vp)], i32),
        "gmpt_forest_destroy": ([vp], None),
        "gmpt_last_error": ([], ctypes.c_char_p),
        "gmpt_puts_2phase": ([vp, ctypes.POINTER(_Batch), ctypes.POINTER(vp), ctypes.POINTER(i32)], i32),
        "gmpt_puts_latching": ([vp, ctypes.POINTER(_Batch), ctypes.POINTER(vp), ctypes.POINTER(i32)], i32),
        "gmpt_puts_with_read": ([vp, ctypes.POINTER(_Batch), vp, i32, ctypes.POINTER(i32), vp, vp,
                                 ctypes.POINTER(vp), ctypes.POINTER(i32)], i32),
        "gmpt_hash_onepass": ([vp, vp, i32], i32),
        "gmpt_get_root_hash": ([vp, u32, vp, ctypes.POINTER(i32)], i32),
        "gmpt_get_root_hashes": ([vp, u32, u32, vp, vp], i32),
        "gmpt_gets": ([vp, ctypes.POINTER(_Batch), vp, vp], i32),
        "gmpt_get_values": ([vp, ctypes.POINTER(_Batch), vp, i64, vp, vp], i32),
        "gmpt_export_nodes": ([vp, ctypes.POINTER(_Batch), ctypes.POINTER(vp), ctypes.POINTER(vp),
                               ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64)], i32),
        "gmpt_free_host": ([vp], None),
        "gmpt_reset_trie": ([vp, u32], i32),
        "gmpt_forest_clear": ([vp], i32),
        "gmpt_synchronize": ([vp], i32),
        "gmpt_set_option": ([vp, ctypes.c_char_p, i64], i32),
        "gmpt_get_option": ([vp, ctypes.c_char_p, ctypes.POINTER(i64)], i32),
        "gmpt_autotune": ([vp, ctypes.POINTER(_Batch), i32, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_double)],
                          i32),
        "gmpt_get_stats": ([vp, ctypes.POINTER(Stats)], i32),
        "gmpt_device_count": ([ctypes.POINTER(i32)], i32),
        "gmpt_device_malloc": ([ctypes.POINTER(vp), ctypes.c_size_t], i32),
        "gmpt_device_free": ([vp], i32),
        "gmpt_memcpy_htod": ([vp, vp, ctypes.c_size_t], i32),
        "gmpt_memcpy_dtoh": ([vp, vp, ctypes.c_size_t], i32),
        "gmpt_device_copy": ([vp, vp, ctypes.c_size_t], i32),
        "gmpt_timer_start": ([vp], i32),
        "gmpt_timer_stop": ([vp, ctypes.POINTER(ctypes.c_float)], i32),
        "gmpt_keccak256_batch": ([vp, vp, vp, u32, vp], i32),
        "gmpt_keccak256_batch_mode": ([vp, vp, vp, u32, vp, i32], i32),
        "gmpt_build_root_noalloc": ([ctypes.POINTER(_Batch), i32, vp, ctypes.POINTER(i32)], i32),
        "gmpt_rebuild_graph": ([vp, ctypes.POINTER(_Batch), ctypes.POINTER(i32)], i32),
        "gmpt_server_create": ([i32, u32, u32, ctypes.c_uint64, u32, i32, ctypes.POINTER(vp)], i32),
        "gmpt_server_destroy": ([vp], None),
        "gmpt_server_submit": ([vp, u32, ctypes.POINTER(_Batch), ctypes.POINTER(ctypes.c_uint64)], i32),
        "gmpt_server_poll": ([vp, ctypes.c_uint64, ctypes.POINTER(i32), vp, ctypes.POINTER(i32),
                              ctypes.POINTER(ctypes.c_double)], i32),
        "gmpt_server_wait": ([vp, ctypes.c_uint64, ctypes.POINTER(i32), vp, ctypes.POINTER(i32),
                              ctypes.POINTER(ctypes.c_double)], i32),
        "gmpt_server_drain": ([vp], i32),
        "gmpt_server_get_stats": ([vp, ctypes.POINTER(ServerStats)], i32),
        "gmpt_server_latencies": ([vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)], i32),
        "build_mpt_2phase": ([i32, vp, vp, vp, vp, vp, i32], vp),
        "build_mpt_olc": ([i32, vp, vp, vp, vp, vp, i32], vp),
        "preprocess": ([], i32),
        "get_all_nodes": ([i32, vp, vp, i32], _NodeSet),
        "init_mpt": ([], vp),
        "insert_mpt_2phase": ([vp, vp, vp, vp, vp, vp, i32], vp),
        "insert_mpt_olc": ([vp, vp, vp, vp, vp, vp, i32], vp),
        "free_mpt": ([vp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise GmptError(lib().gmpt_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def host_batch(b, hps=None):
    """A gmpt_batch struct over a workloads.Batch held in host memory."""
    s = _Batch()
    s.keys_hexs, s.keys_indexs = _p(b.keys_hexs), _p(b.keys_indexs)
    s.values_bytes, s.values_indexs = _p(b.values_bytes), _p(b.values_indexs)
    s.values_hps = _p(hps)
    s.trie_ids = _p(b.trie_ids)
    s.n = b.n
    s._keep = (b, hps)
    return s


class DeviceBuffer:
    """A device allocation made through the library (no other runtime needed)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().gmpt_device_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            _check(lib().gmpt_memcpy_htod(d.ptr, _p(a), a.nbytes))
        return d

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib().gmpt_memcpy_dtoh(_p(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().gmpt_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBatch:
    """A workloads.Batch uploaded to HBM; .struct is the device gmpt_batch."""

    def __init__(self, b, adopt_values=False):
        self.bufs = [DeviceBuffer.from_array(x) for x in (b.keys_hexs, b.keys_indexs, b.values_bytes,
                                                         b.values_indexs)]
        self.tid = DeviceBuffer.from_array(b.trie_ids) if b.trie_ids is not None else None
        s = _Batch()
        s.keys_hexs, s.keys_indexs, s.values_bytes, s.values_indexs = [x.ptr for x in self.bufs]
        s.values_hps = None
        s.trie_ids = self.tid.ptr if self.tid else None
        s.n = b.n
        s.keys_bytes = b.key_bytes
        s.values_bytes_total = b.value_bytes
        s.on_device = 1
        s.adopt_values = 1 if adopt_values else 0
        self.struct = s
        self.n = b.n
        self.holders = 0  # tries (MPT objects) that point into the value buffer

    def view(self, start, count):
        """Keys [start, start + count) as a single-trie device batch sharing
        this batch's buffers (index pairs keep addressing the whole buffers)."""
        v = DeviceBatch.__new__(DeviceBatch)
        s = _Batch.from_buffer_copy(self.struct)
        s.keys_indexs = self.struct.keys_indexs + 8 * start
        s.values_indexs = self.struct.values_indexs + 16 * start
        s.trie_ids = None
        s.n = count
        v.bufs, v.tid, v.struct, v.n, v.holders = [], None, s, count, 0
        v._parent = self  # keeps the shared buffers alive
        return v

    def free(self):
        """Releases the device buffers; refused while a trie still points into
        the value buffer (adopt_values) -- clear() or close() that trie first."""
        if self.holders:
            raise GmptError("DeviceBatch.free: a trie still refers to this batch's values "
                            "(clear() or close() it first)")
        for x in self.bufs:
            x.free()
        if self.tid:
            self.tid.free()


def build_root_noalloc(batch, mode=0):
    """Root hash of the trie `batch` builds from empty, computed by the
    library's no-allocation entry point (a fixed device region per GPU, the
    current device, at most NOALLOC_MAX_KEYS keys).  mode 0 PhaseNU, 1 LockNU."""
    s = batch.struct if isinstance(batch, DeviceBatch) else host_batch(batch)
    out = np.zeros(32, np.uint8)
    sz = ctypes.c_int()
    _check(lib().gmpt_build_root_noalloc(ctypes.byref(s), mode, _p(out), ctypes.byref(sz)))
    return out[:sz.value].tobytes()


NOALLOC_MAX_KEYS = 32768


class MPT:
    """GpuMPT::Compress::MPT over a forest of `num_tries` tries on one device."""

    def __init__(self, device=0, num_tries=1):
        self._f = ctypes.c_void_p()
        _check(lib().gmpt_forest_create(device, num_tries, ctypes.byref(self._f)))
        self.num_tries = num_tries
        # device batches whose value buffers the tries point into
        # (adopt_values): kept alive until the tries stop referring to them
        self._adopted = []

    def close(self):
        if self._f and self._f.value:
            lib().gmpt_forest_destroy(self._f)
            self._f = None
        if getattr(self, "_adopted", None):
            self._release_adopted()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- inserts (gpu_mpt.cuh:69-71, 95-97) -----------------------------------
    def _adopt(self, batch):
        if isinstance(batch, DeviceBatch) and batch.struct.adopt_values and batch.n:
            owner = getattr(batch, "_parent", batch)  # a view's buffers belong to its parent
            owner.holders += 1
            self._adopted.append(owner)

    def _puts(self, fn, batch, hps=None):
        s = batch.struct if isinstance(batch, DeviceBatch) else host_batch(batch, hps)
        t = ctypes.c_void_p()
        n = ctypes.c_int()
        _check(fn(self._f, ctypes.byref(s), ctypes.byref(t), ctypes.byref(n)))
        self._adopt(batch)
        return t.value, n.value

    def puts_2phase_with_valuehp(self, batch, values_hps=None):
        """PhaseNU: returns (hash targets, count) for hash_onepass_v2."""
        return self._puts(lib().gmpt_puts_2phase, batch, values_hps)

    def puts_latching_with_valuehp_v2(self, batch, values_hps=None):
        """LockNU: returns (hash targets, count) for hash_onepass_v2."""
        return self._puts(lib().gmpt_puts_latching, batch, values_hps)

    def puts_with_read(self, batch, rw_flags, mode=0, values_hps=None):
        """Mixed batch (gpu_mpt.cuh:74-78, 105-110): rw_flags[i] 0 = read, 1 =
        write.  Returns (targets, count, reads) with reads = [(host pointer,
        size)] in batch order of the reads, seen before the batch's writes."""
        s = batch.struct if isinstance(batch, DeviceBatch) else host_batch(batch, values_hps)
        flags = rw_flags if isinstance(rw_flags, DeviceBuffer) else np.ascontiguousarray(rw_flags, np.uint8)
        fptr = flags.ptr if isinstance(flags, DeviceBuffer) else _p(flags)
        n = batch.n
        rn = ctypes.c_int()
        hps = np.zeros(max(n, 1), np.uint64)
        sizes = np.zeros(max(n, 1), np.int32)
        t = ctypes.c_void_p()
        nt = ctypes.c_int()
        _check(lib().gmpt_puts_with_read(self._f, ctypes.byref(s), fptr, mode, ctypes.byref(rn), _p(hps), _p(sizes),
                                         ctypes.byref(t), ctypes.byref(nt)))
        self._adopt(batch)
        reads = [(int(hps[i]), int(sizes[i])) for i in range(rn.value)]
        return t, nt.value, reads

    def rebuild_graph(self, batch):
        """Empties the forest and bulk-builds `batch` (a DeviceBatch made with
        adopt_values=True) as one replay of a captured HIP graph; the first
        call for a batch's buffers captures.  Returns True when a replay stood
        (False: the pass ran eagerly, and was captured for the next call)."""
        if not isinstance(batch, DeviceBatch) or not batch.struct.adopt_values:
            raise GmptError("rebuild_graph takes a DeviceBatch with adopt_values=True")
        rep = ctypes.c_int()
        self._release_adopted()  # the clear drops the trie's references
        _check(lib().gmpt_rebuild_graph(self._f, ctypes.byref(batch.struct), ctypes.byref(rep)))
        self._adopt(batch)
        return bool(rep.value)

    # --- hashing (gpu_mpt.cuh:131) --------------------------------------------
    def hash_onepass_v2(self, targets, n):
        _check(lib().gmpt_hash_onepass(self._f, targets, n))

    def get_root_hash(self, trie=0):
        """Root hash bytes (32, or the raw root encoding when < 32 B, b'' if empty)."""
        out = np.zeros(32, np.uint8)
        sz = ctypes.c_int()
        _check(lib().gmpt_get_root_hash(self._f, trie, _p(out), ctypes.byref(sz)))
        return out[:sz.value].tobytes()

    def get_root_hashes(self, first=0, count=None):
        """Root hashes of tries [first, first + count) read with one copy."""
        count = self.num_tries - first if count is None else count
        out = np.zeros(32 * max(count, 1), np.uint8)
        sz = np.zeros(max(count, 1), np.int32)
        _check(lib().gmpt_get_root_hashes(self._f, first, count, _p(out), _p(sz)))
        return [out[32 * i:32 * i + sz[i]].tobytes() for i in range(count)]

    # --- gets (gpu_mpt.cuh:139-140) -------------------------------------------
    def gets_parallel(self, batch):
        """(host pointers recorded at insert, sizes); absent keys give (0, 0)."""
        n = batch.n
        hps = np.zeros(max(n, 1), np.uint64)
        sizes = np.zeros(max(n, 1), np.int32)
        _check(lib().gmpt_gets(self._f, ctypes.byref(host_batch(batch)), _p(hps), _p(sizes)))
        return hps[:n], sizes[:n]

    def get_values(self, batch):
        """Value bytes per key (None when absent), read from HBM."""
        n = batch.n
        offs = np.zeros(max(n, 1), np.int64)
        sizes = np.zeros(max(n, 1), np.int32)
        cap = 1 << 20
        while True:
            out = np.zeros(cap, np.uint8)
            rc = lib().gmpt_get_values(self._f, ctypes.byref(host_batch(batch)), _p(out), cap, _p(offs), _p(sizes))
            if rc == 0:
                break
            msg = lib().gmpt_last_error().decode()
            if "too small" not in msg:
                raise GmptError(msg)
            cap *= 4
        return [None if sizes[i] < 0 else out[offs[i]:offs[i] + sizes[i]].tobytes() for i in range(n)]

    # --- node export (gpu_mpt.cuh:146-149) ------------------------------------
    def flush_dirty_nodes(self, batch):
        """[(32-byte hash slot, RLP encoding)] of every node on the keys' paths."""
        h, e, ix = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        num = ctypes.c_uint64()
        _check(lib().gmpt_export_nodes(self._f, ctypes.byref(host_batch(batch)), ctypes.byref(h), ctypes.byref(e),
                                       ctypes.byref(ix), ctypes.byref(num)))
        n = num.value
        out = []
        if n:
            hashes = np.ctypeslib.as_array(ctypes.cast(h, ctypes.POINTER(ctypes.c_uint8)), shape=(32 * n,)).copy()
            idx = np.ctypeslib.as_array(ctypes.cast(ix, ctypes.POINTER(ctypes.c_uint64)), shape=(2 * n,)).copy()
            tot = int(idx[-1]) + 1
            encs = np.ctypeslib.as_array(ctypes.cast(e, ctypes.POINTER(ctypes.c_uint8)), shape=(tot,)).copy()
            for i in range(n):
                out.append((hashes[32 * i:32 * i + 32].tobytes(), encs[idx[2 * i]:idx[2 * i + 1] + 1].tobytes()))
        for p in (h, e, ix):
            if p.value:
                lib().gmpt_free_host(p)
        return out

    def reset_trie(self, trie):
        _check(lib().gmpt_reset_trie(self._f, trie))

    def _release_adopted(self):
        for b in self._adopted:
            b.holders -= 1
        self._adopted = []

    def clear(self):
        _check(lib().gmpt_forest_clear(self._f))
        self._release_adopted()

    def synchronize(self):
        _check(lib().gmpt_synchronize(self._f))

    def set_option(self, name, value):
        _check(lib().gmpt_set_option(self._f, name.encode(), int(value)))

    def get_option(self, name):
        v = ctypes.c_int64()
        _check(lib().gmpt_get_option(self._f, name.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, sample, mode=0):
        """Picks hc_spread_max by timing PhaseHC on `sample`; returns (value, ms)."""
        s = sample.struct if isinstance(sample, DeviceBatch) else host_batch(sample)
        v = ctypes.c_int64()
        ms = ctypes.c_double()
        _check(lib().gmpt_autotune(self._f, ctypes.byref(s), mode, ctypes.byref(v), ctypes.byref(ms)))
        return v.value, ms.value

    def stats(self):
        s = Stats()
        _check(lib().gmpt_get_stats(self._f, ctypes.byref(s)))
        return s.as_dict()

    def timer_start(self):
        _check(lib().gmpt_timer_start(self._f))

    def timer_stop(self):
        ms = ctypes.c_float()
        _check(lib().gmpt_timer_stop(self._f, ctypes.byref(ms)))
        return ms.value


class Server:
    """Continuous-batching insert + hash service over `num_tries` HBM-resident
    tries (gmpt_server_* in gmpt_amd.h)."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, -1

    def __init__(self, device=0, num_tries=1, lanes=2, mem_budget=0, max_batch_keys=0, mode=0):
        self._s = ctypes.c_void_p()
        _check(lib().gmpt_server_create(device, num_tries, lanes, mem_budget, max_batch_keys, mode,
                                        ctypes.byref(self._s)))
        self._keep = {}

    def submit(self, trie, batch):
        """Queues one host batch for `trie`; returns the request id."""
        s = host_batch(batch)
        rid = ctypes.c_uint64()
        _check(lib().gmpt_server_submit(self._s, trie, ctypes.byref(s), ctypes.byref(rid)))
        return rid.value

    def _result(self, fn, rid):
        st = ctypes.c_int()
        root = np.zeros(32, np.uint8)
        sz = ctypes.c_int()
        lat = ctypes.c_double()
        _check(fn(self._s, rid, ctypes.byref(st), _p(root), ctypes.byref(sz), ctypes.byref(lat)))
        if st.value == self.FAILED:
            raise GmptError(lib().gmpt_last_error().decode())
        if st.value != self.DONE:
            return st.value, None, None
        return st.value, root[:sz.value].tobytes(), lat.value

    def poll(self, rid):
        """(state, root or None, latency ms or None)."""
        return self._result(lib().gmpt_server_poll, rid)

    def wait(self, rid):
        """Blocks until request `rid` completes: (root, latency ms)."""
        _, root, lat = self._result(lib().gmpt_server_wait, rid)
        return root, lat

    def drain(self):
        _check(lib().gmpt_server_drain(self._s))

    def stats(self):
        s = ServerStats()
        _check(lib().gmpt_server_get_stats(self._s, ctypes.byref(s)))
        return s.as_dict()

    def latencies(self):
        cnt = ctypes.c_uint64()
        _check(lib().gmpt_server_latencies(self._s, None, 0, ctypes.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), np.float64)
        _check(lib().gmpt_server_latencies(self._s, _p(out), cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    def close(self):
        if self._s and self._s.value:
            lib().gmpt_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def keccak256_batch(messages, mode=0):
    """Keccak-256 of each message, computed on the GPU (mode 0: one state per
    lane, mode 1: one state spread over 25 lanes)."""
    n = len(messages)
    lens = np.array([len(m) for m in messages], np.uint32)
    offs = np.zeros(n, np.uint64)
    if n:
        offs[1:] = np.cumsum(lens[:-1].astype(np.uint64))
    data = np.frombuffer(b"".join(messages) + b"\0" * 8, np.uint8).copy()
    out = np.zeros(32 * max(n, 1), np.uint8)
    _check(lib().gmpt_keccak256_batch_mode(_p(data), _p(offs), _p(lens), n, _p(out), mode))
    return [out[32 * i:32 * i + 32].tobytes() for i in range(n)]


def device_count():
    c = ctypes.c_int()
    _check(lib().gmpt_device_count(ctypes.byref(c)))
    return c.value
