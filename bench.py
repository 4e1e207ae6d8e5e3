"""Benchmark: MPT puts + root hash on one MI355X.

Default workload (the configuration the headline metric is quoted on): one
batch of 1,000,000 unique uniform 32-byte keys with 70..110-byte random values
inserted into an empty trie with PhaseNU, then the one-pass root hash
(PhaseHC).  Inputs are resident in HBM before the timed region (the trie
adopts the value buffer); each step clears the forest, inserts, hashes and
reads the 32-byte root.  The root of every step is checked against the
oracle's root of the same batch.

Prints ONE JSON line (rank 0).  Other workloads: --config lock (LockNU under
Zipf skew), --config batched (4,096 uneven tries), --config served.
"""
import argparse
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "gpu-merkle-patricia-trie_amd")

# MI355X published peaks used for the roofline
HBM_PEAK_GBS = 8000.0
CUS, LANES, PEAK_CLOCK_GHZ = 256, 64, 2.4
VALU_PEAK_TOPS = CUS * LANES * PEAK_CLOCK_GHZ * 1e9 / 1e12  # 32-bit lane operations
# 64-bit logical ops per second when every 64-bit xor/and/rotate costs two
# 32-bit lane-ops
VALU64_PEAK_TOPS = VALU_PEAK_TOPS / 2
OPS64_PER_PERM = 3720  # Keccak-f[1600]: 155 64-bit lane-ops per round x 24 rounds
# The fewest 32-bit VALU instructions a Keccak-f[1600] needs on gfx950, whose
# 3-input v_bitop3_b32 does xor3 and chi's a ^ (~b & c) in one op: per round
# theta parities 20, parity rotations 10, theta updates 50, rho 48, chi 50,
# iota 2 = 180; x 24 rounds.
INSTR_PER_PERM = 4320
PERM_PEAK_G = VALU_PEAK_TOPS * 1e12 / INSTR_PER_PERM / 1e9


# HBM bytes per launch of the leaf-level kernel on the default 1M-key
# workload: rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) per dispatch,
# the factors being what the same counters read for the 1 GiB k_copy16
# calibration copy (--config hbm-copy).  Keyed by "leaves fused".
LEAF_TRAFFIC_BYTES = {False: None, True: 586_695_130}


def _import_package():
    spec = importlib.util.spec_from_file_location("gmpt_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmpt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


class _OneCore:
    """Pins the process to one core for a CPU baseline, restoring the
    previous affinity afterwards (the serving runs use host threads)."""

    def __enter__(self):
        self.prev = None
        try:
            self.prev = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {sorted(self.prev)[0]})
        except (AttributeError, OSError):
            pass
        return self

    def __exit__(self, *exc):
        if self.prev:
            try:
                os.sched_setaffinity(0, self.prev)
            except OSError:
                pass
        return False


def cpu_baseline_build(batch, label, runs=5):
    """The oracle (C restatement of the reference CPU MPT) timed on one core:
    one warm-up build, then the median of `runs` timed builds (insert + hash
    in one timed region, as the reference's CPU baseline times it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    secs = []
    with _OneCore():
        _, root = oracle.time_build(batch)  # warm-up
        for _ in range(runs):
            dt, r = oracle.time_build(batch)
            if r != root:
                raise SystemExit("bench: oracle roots differ between runs")
            secs.append(dt)
    med = float(np.median(secs))
    return {"value": batch.n / med, "unit": "puts+root-hash/s", "cores": 1, "kind": "port",
            "sample": f"{label}: puts_baseline + hashs_dirty_flag, 1 warm-up + median of {runs} runs "
                      f"({med:.2f} s each, {min(secs):.2f}-{max(secs):.2f} s)",
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}, root


def bench_phase(pkg, args):
    wl = pkg.workloads
    t0 = time.time()
    b = wl.config2_phase(n=args.n)
    gen_s = time.time() - t0
    db = pkg.DeviceBatch(b, adopt_values=True)
    t = pkg.MPT(device=0)
    t.set_option("hc_fuse_leaves", 0 if args.no_fuse_leaves else 1)
    t.set_option("hc_fuse_inner", 0 if args.no_fuse_inner else 1)
    if args.pool_host:  # the trie state in pinned host memory (tries beyond HBM)
        t.set_option("pool_memory", 1)
    fused = not args.no_fuse_leaves
    if args.spread_max is None:
        spread_max, tune_ms = t.autotune(db)  # picks the Keccak layout threshold on this batch
    else:  # fixed threshold: no autotuner passes (clean kernel traces)
        t.set_option("hc_spread_max", args.spread_max)
        spread_max, tune_ms = args.spread_max, None
    roots = []

    def step():
        t.clear()
        targets, nt = t.puts_2phase_with_valuehp(db)
        t.hash_onepass_v2(targets, nt)
        roots.append(t.get_root_hash())

    for _ in range(args.warmup):
        step()
    st = t.stats()
    t.synchronize()
    hash_ms, leaf_ms_l = [], []
    t.timer_start()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        s1 = t.stats()
        hash_ms.append(s1["last_hash_ms"])
        leaf_ms_l.append(s1["leaf_perm_ms"])
    t.synchronize()
    wall = time.perf_counter() - w0
    dev_ms = t.timer_stop()
    st = t.stats()
    ms_per_step = wall * 1e3 / args.steps
    value = b.n * args.steps / wall

    # roofline of the dominant kernel: the leaf level's Keccak permutations
    # (k_hc_leaf_direct, which also encodes the leaves, or k_hc_perm_lane),
    # timed with HIP events on the forest's stream
    perms = st["perms"]
    hms = sum(hash_ms) / len(hash_ms)
    leaf_ms = sum(leaf_ms_l) / len(leaf_ms_l)
    achieved = st["leaf_perms"] / (leaf_ms * 1e-3) / 1e9
    # whole-path roofline: input bytes + node bytes over HBM, Keccak over VALU
    b_in = b.key_bytes + 8 * b.n + b.value_bytes + 16 * b.n + 8 * b.n
    b_nodes = st["enc_bytes"] + 32 * st["hashed_nodes"]
    t_hbm = (b_in + b_nodes) / (HBM_PEAK_GBS * 1e9)
    t_roof = t_hbm + OPS64_PER_PERM * perms / (VALU64_PEAK_TOPS * 1e12)
    t_roof_isa = t_hbm + perms / (PERM_PEAK_G * 1e9)

    out = {
        "metric": "MPT puts+root-hash ops/s at 1M 32B keys", "value": value, "unit": "puts+root-hash/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8/u64 (integer)",
        "data": "synthetic (seeded uniform 32-byte keys, 70-110-byte random values)",
        "config": {"workload": "PhaseNU+PhaseHC, 1,000,000 unique 32-byte keys, 70-110-byte values, "
                               "one batch into an empty trie, inputs resident in HBM",
                   "keys": b.n, "value_bytes": b.value_bytes,
                   "trie_state": "pinned host memory" if args.pool_host else "HBM"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PERM_PEAK_G, "unit": "Gperm/s",
                     "frac": achieved / PERM_PEAK_G,
                     "traffic": LEAF_TRAFFIC_BYTES[fused] if b.n == 1_000_000 else None,
                     "kernel": ("k_hc_leaf_direct, leaf level (leaf encoding built word by word in registers "
                                "+ Keccak-f[1600], one state per lane)") if fused else
                               "k_hc_perm_lane, leaf level (Keccak-f[1600], one state per lane)",
                     "per_launch": {"permutations": st["leaf_perms"], "nodes": st["leaf_nodes"],
                                    "ms": leaf_ms, "message_bytes": st["leaf_bytes"]},
                     "peak_basis": f"{VALU_PEAK_TOPS:.2f} T 32-bit VALU ops/s / {INSTR_PER_PERM} ops per permutation"},
        "hash": {"permutations": perms, "nodes_hashed": st["hashed_nodes"], "hash_ms": hms,
                 "launches": st["hash_launches"], "gperm_per_s": perms / (hms * 1e-3) / 1e9},
        "path_roofline": {"t_roof_ms": t_roof * 1e3, "t_measured_ms": ms_per_step,
                          "frac": t_roof * 1e3 / ms_per_step, "t_roof_isa_ms": t_roof_isa * 1e3,
                          "frac_isa": t_roof_isa * 1e3 / ms_per_step, "bytes_in": b_in, "bytes_nodes": b_nodes,
                          "device_ms_per_step": dev_ms / args.steps},
        "mputs_per_s": value / 1e6, "mhashes_per_s": st["hashed_nodes"] / (ms_per_step * 1e-3) / 1e6,
        "autotune": {"hc_spread_max": spread_max, "hash_ms": tune_ms},
        "root": roots[-1].hex(),
    }
    if len(set(roots)) != 1:
        raise SystemExit(f"bench: roots differ between steps: {set(r.hex() for r in roots)}")
    # the same step from host buffers: the library stages keys, indexes and
    # values over PCIe inside the timed region (reported, never `value`)
    host_s = []
    for k in range(6):
        t.clear()
        t.synchronize()
        w1 = time.perf_counter()
        targets, nt = t.puts_2phase_with_valuehp(b)
        t.hash_onepass_v2(targets, nt)
        hroot = t.get_root_hash()
        if k:
            host_s.append(time.perf_counter() - w1)
        if hroot != roots[-1]:
            raise SystemExit("bench: host-batch root differs")
    out["pcie_inclusive"] = {"puts_per_s": b.n / (sum(host_s) / len(host_s)),
                             "ms_per_step": 1e3 * sum(host_s) / len(host_s),
                             "bytes_staged": int(b_in - 8 * b.n)}
    if not args.no_cpu_baseline:
        cb, croot = cpu_baseline_build(b, "1,000,000 keys (the same batch)")
        if croot != roots[-1]:
            raise SystemExit(f"bench: GPU root {roots[-1].hex()} != oracle root {croot.hex()}")
        out["cpu_baseline"] = cb
        out["root_matches_oracle"] = True
    else:
        out["cpu_baseline"] = None
    out["gen_s"] = gen_s
    t.close()
    db.free()
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def bench_batched(pkg, args):
    """4,096 uneven tries: one multi-trie PhaseNU insert + one PhaseHC pass for
    all of them, against one plain insert + hash + root call per trie."""
    wl = pkg.workloads
    t0 = time.time()
    b, sizes = wl.config4_batched()
    gen_s = time.time() - t0
    ntries = len(sizes)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    db = pkg.DeviceBatch(b, adopt_values=True)
    views = [db.view(int(s), int(c)) for s, c in zip(starts, sizes)]
    fb = pkg.MPT(device=0, num_tries=ntries)
    fp = pkg.MPT(device=0)

    def batched_step():
        fb.clear()
        targets, nt = fb.puts_2phase_with_valuehp(db)
        fb.hash_onepass_v2(targets, nt)
        return fb.get_root_hashes()

    def plain_step():
        roots = []
        for v in views:
            fp.clear()
            targets, nt = fp.puts_2phase_with_valuehp(v)
            fp.hash_onepass_v2(targets, nt)
            roots.append(fp.get_root_hash())
        return roots

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            out = fn()
        fb.synchronize()
        fp.synchronize()
        w0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        fb.synchronize()
        fp.synchronize()
        return (time.perf_counter() - w0) / steps, out

    bt, broots = timed(batched_step, args.steps, args.warmup)
    st = fb.stats()
    # the same pass as a replayed HIP graph (the first call captures)
    fg = pkg.MPT(device=0, num_tries=ntries)
    replays = []

    def graph_step():
        replays.append(fg.rebuild_graph(db))
        return fg.get_root_hashes()

    fg.rebuild_graph(db)  # capture
    gt, groots = timed(graph_step, args.steps, args.warmup)
    if groots != broots or not all(replays):
        raise SystemExit("bench: graph replay roots differ from the eager pass, or a replay did not stand")
    pt, proots = timed(plain_step, max(1, args.steps // 5), 1)
    if broots != proots:
        bad = [i for i in range(ntries) if broots[i] != proots[i]][:5]
        raise SystemExit(f"bench: batched and per-trie roots differ for tries {bad}")
    oracle = _oracle()
    rng = np.random.default_rng(0)
    sample = sorted(set(rng.choice(ntries, 48, replace=False).tolist()) | {int(np.argmax(sizes))})
    for i in sample:
        want = oracle.root_of(b.subset(range(int(starts[i]), int(starts[i] + sizes[i]))))
        if broots[i] != want:
            raise SystemExit(f"bench: trie {i} root {broots[i].hex()} != oracle {want.hex()}")
    out = {
        "metric": "batched tries/s (4,096 uneven tries, insert + root hash)", "value": ntries / bt,
        "unit": "tries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": bt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8/u64 (integer)",
        "data": "synthetic (seeded lognormal trie sizes, uniform 32-byte keys, 32-byte random values)",
        "config": {"workload": "4,096 tries, sizes lognormal(median 256, sigma 2.046) clipped to 1..200,000, "
                               "one multi-trie PhaseNU batch + one PhaseHC pass, inputs resident in HBM",
                   "tries": ntries, "keys": int(b.n)},
        "plain": {"tries_per_s": ntries / pt, "ms_per_pass": pt * 1e3,
                  "how": "one insert + hash + root call per trie on a reused single-trie forest"},
        "graph": {"tries_per_s": ntries / gt, "ms_per_pass": gt * 1e3, "replays_stood": len(replays),
                  "how": "the same multi-trie pass replayed as one captured HIP graph (gmpt_rebuild_graph)"},
        "speedup_vs_plain": pt / bt,
        "mputs_per_s": b.n / bt / 1e6,
        "per_step": {"permutations": st["perms"], "nodes_hashed": st["hashed_nodes"],
                     "hash_ms": st["last_hash_ms"], "launches": st["hash_launches"]},
        "roots_checked_vs_oracle": len(sample), "gen_s": gen_s,
    }
    if not args.no_cpu_baseline:
        # sequential oracle build + hash of whole tries, on one core, a
        # bounded sample (cpu_seconds of builds)
        order = rng.permutation(ntries)
        done, keys, secs = 0, 0, 0.0
        with _OneCore():
            for i in order:
                sub = b.subset(range(int(starts[i]), int(starts[i] + sizes[i])))
                dt, _ = oracle.time_build(sub)
                secs += dt
                done += 1
                keys += int(sizes[i])
                if secs > args.cpu_seconds:
                    break
        out["cpu_baseline"] = {"value": done / secs, "unit": "tries/s", "cores": 1, "kind": "port",
                               "sample": f"{done} random whole tries ({keys} keys), puts + hash each, {secs:.1f} s",
                               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    else:
        out["cpu_baseline"] = None
    fb.close()
    fp.close()
    fg.close()
    db.free()
    return out


def bench_lock(pkg, args):
    """LockNU under Zipf(0.99) skew: 4M preloaded keys, then 1M overwrite/insert
    ops with 1,000-byte values; the timed region is the ops' insert + hash."""
    wl = pkg.workloads
    t0 = time.time()
    pre, ops = wl.config3_lock()
    gen_s = time.time() - t0
    dpre = pkg.DeviceBatch(pre, adopt_values=True)
    dops = pkg.DeviceBatch(ops, adopt_values=True)
    t = pkg.MPT(device=0)
    roots, times, stats = [], [], []
    for k in range(args.warmup + args.steps):
        t.clear()
        targets, nt = t.puts_2phase_with_valuehp(dpre)
        t.hash_onepass_v2(targets, nt)
        t.synchronize()
        w0 = time.perf_counter()
        targets, nt = t.puts_latching_with_valuehp_v2(dops)
        s_ins = t.stats()
        t.hash_onepass_v2(targets, nt)
        roots.append(t.get_root_hash())
        dt = time.perf_counter() - w0
        if k >= args.warmup:
            times.append(dt)
            st = t.stats()
            st["restarts"], st["lock_fails"] = s_ins["restarts"], s_ins["lock_fails"]
            stats.append(st)
    if len(set(roots)) != 1:
        raise SystemExit("bench: LockNU roots differ between steps")
    sec = float(np.mean(times))
    out = {
        "metric": "LockNU puts+root-hash ops/s (Zipf 0.99, 50% overwrite)", "value": ops.n / sec,
        "unit": "puts+root-hash/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u64 (integer)",
        "data": "synthetic ('user'+19-digit Zipf ids over 10M, 1,000-byte random values)",
        "config": {"workload": "4M preloaded keys (untimed), then 1M LockNU ops, 50% overwrite / 50% insert, "
                               "Zipf(0.99) ids, 1,000-byte values, then one-pass root hash",
                   "preload": pre.n, "ops": ops.n},
        "per_step": {"restarts": stats[-1]["restarts"], "lock_fails": stats[-1]["lock_fails"],
                     "permutations": stats[-1]["perms"], "nodes_hashed": stats[-1]["hashed_nodes"],
                     "hash_ms": stats[-1]["last_hash_ms"]},
        "root": roots[-1].hex(), "gen_s": gen_s,
    }
    if not args.no_cpu_baseline:
        oracle = _oracle()
        o = oracle.OracleMPT()
        with _OneCore():
            o.puts(pre)
            o.hash()
            w0 = time.perf_counter()
            o.puts(ops)
            o.hash()
            secs = time.perf_counter() - w0
        if o.root() != roots[-1]:
            raise SystemExit(f"bench: GPU root {roots[-1].hex()} != oracle root {o.root().hex()}")
        out["root_matches_oracle"] = True
        out["cpu_baseline"] = {"value": ops.n / secs, "unit": "puts+root-hash/s", "cores": 1, "kind": "port",
                               "sample": f"the full 1M ops applied sequentially to the preloaded trie, {secs:.1f} s",
                               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    else:
        out["cpu_baseline"] = None
    t.close()
    dpre.free()
    dops.free()
    return out


def _poisson_run(submit, reqs, rate, duration, rng):
    """Submits reqs cyclically at Poisson arrivals of `rate` per second for
    `duration` seconds; returns the number submitted."""
    t0 = time.perf_counter()
    t_next = t0
    k = 0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next - t0 > duration:
            return k
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        submit(*reqs[k % len(reqs)])
        k += 1


def _pct(lat):
    lat = np.asarray(lat, np.float64)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "requests": int(lat.size)}


def bench_served(pkg, args):
    """Continuous batching: Poisson arrivals of insert+hash requests over 1,024
    resident tries (Zipf 0.8), against one plain insert+hash call per request."""
    import queue
    import threading
    wl = pkg.workloads
    gmpt = pkg.gmpt
    tries = 1024
    t0 = time.time()
    pool = wl.served_requests(n_requests=args.pool)
    pre = wl.served_requests(seed=55, n_requests=tries)
    gen_s = time.time() - t0
    rng = np.random.default_rng(5)
    srv = gmpt.Server(device=0, num_tries=tries, lanes=args.lanes, max_batch_keys=args.max_batch_keys)
    for t in range(tries):  # resident tries, pre-populated (untimed)
        srv.submit(t, pre[t][1])
    srv.drain()

    # saturation: a backlog of requests submitted at once
    n0 = srv.stats()["completed"]
    w0 = time.perf_counter()
    for t, b in pool[:args.backlog]:
        srv.submit(t, b)
    srv.drain()
    sat_s = time.perf_counter() - w0
    sat_rps = args.backlog / sat_s
    sat_keys = sum(b.n for _, b in pool[:args.backlog])

    # offered load: Poisson arrivals at a fraction of saturation
    runs = []
    for frac in args.rates:
        base = len(srv.latencies())
        st0 = srv.stats()
        w0 = time.perf_counter()
        k = _poisson_run(lambda t, b: srv.submit(t, b), pool, frac * sat_rps, args.duration, rng)
        srv.drain()
        wall = time.perf_counter() - w0
        st1 = srv.stats()
        lat = srv.latencies()[base:]
        keys = st1["batched_keys"] - st0["batched_keys"]
        batches = st1["batches"] - st0["batches"]
        runs.append({"offered_frac_of_saturation": frac, "offered_rps": frac * sat_rps, "requests": k,
                     "puts_per_s": keys / wall, "requests_per_s": k / wall,
                     "requests_per_batch": (st1["batched_requests"] - st0["batched_requests"]) / max(batches, 1),
                     "deferred_by_memory": st1["deferred_mem"] - st0["deferred_mem"], **_pct(lat)})
    sst = srv.stats()
    srv.close()

    # plain: one insert + hash + root call per request, FIFO, one worker
    f = gmpt.MPT(device=0, num_tries=tries)
    for t in range(tries):
        b = pre[t][1]
        b.trie_ids = np.full(b.n, t, np.int32)
        tg, n = f.puts_2phase_with_valuehp(b)
        f.hash_onepass_v2(tg, n)
    for t, b in pool:
        b.trie_ids = np.full(b.n, t, np.int32)

    def plain_one(t, b):
        tg, n = f.puts_2phase_with_valuehp(b)
        f.hash_onepass_v2(tg, n)
        return f.get_root_hash(t)

    w0 = time.perf_counter()
    for t, b in pool[:args.backlog]:
        plain_one(t, b)
    plain_sat_s = time.perf_counter() - w0
    plain_rps = args.backlog / plain_sat_s
    q = queue.Queue()
    lat = []

    def worker():
        while True:
            item = q.get()
            if item is None:
                return
            ta, t, b = item
            plain_one(t, b)
            lat.append((time.perf_counter() - ta) * 1e3)

    th = threading.Thread(target=worker)
    th.start()
    prate = args.rates[0] * plain_rps
    w0 = time.perf_counter()
    k = _poisson_run(lambda t, b: q.put((time.perf_counter(), t, b)), pool, prate, args.plain_duration, rng)
    q.put(None)
    th.join()
    pwall = time.perf_counter() - w0
    f.close()
    main_run = runs[0]
    out = {
        "metric": "served puts/s with p50/p99 latency (continuous batching, 1,024 resident tries)",
        "value": main_run["puts_per_s"], "unit": "puts/s", "n_gpus": 1, "steps": len(runs), "warmup": 0,
        "ms_per_step": args.duration * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u64 (integer)",
        "data": "synthetic (Zipf(0.8) trie choice, lognormal(median 2,000) request sizes, uniform 32-byte keys, "
                "70-110-byte values)",
        "config": {"workload": f"Poisson arrivals at {args.rates} x measured saturation for {args.duration:.0f} s "
                               "each, 1,024 HBM-resident tries, PhaseNU insert + PhaseHC hash per batch",
                   "lanes": args.lanes, "max_batch_keys": args.max_batch_keys, "request_pool": len(pool)},
        "latency": {k2: main_run[k2] for k2 in ("p50_ms", "p99_ms", "mean_ms")},
        "saturation": {"requests_per_s": sat_rps, "puts_per_s": sat_keys / sat_s},
        "runs": runs,
        "plain": {"saturation_requests_per_s": plain_rps, "offered_rps": prate, "requests": k,
                  "requests_per_s": k / pwall, **_pct(lat),
                  "how": "one insert + hash + root call per request, FIFO, one worker thread"},
        "speedup_vs_plain_saturation": sat_rps / plain_rps,
        "server_stats": sst, "gen_s": gen_s, "cpu_baseline": None,
    }
    return out


def bench_hbm_copy(pkg, args):
    """Calibration: a plain 16-byte-per-lane copy of 1 GiB, the known-traffic
    kernel that scales the FETCH_SIZE / WRITE_SIZE counters (k_copy16)."""
    gmpt = pkg.gmpt
    nbytes = 1 << 30
    src, dst = gmpt.DeviceBuffer(nbytes), gmpt.DeviceBuffer(nbytes)
    for _ in range(args.warmup):
        gmpt.lib().gmpt_device_copy(dst.ptr, src.ptr, nbytes)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        if gmpt.lib().gmpt_device_copy(dst.ptr, src.ptr, nbytes) != 0:
            raise SystemExit("bench: copy failed")
    sec = (time.perf_counter() - w0) / args.steps
    src.free()
    dst.free()
    return {"metric": "HBM copy GB/s (read + write)", "value": 2 * nbytes / sec / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "uninitialised device buffers", "config": {"workload": "k_copy16, 1 GiB -> 1 GiB"},
            "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes, "cpu_baseline": None}


def _sub_args(args, **kw):
    d = dict(vars(args))
    d.update(kw)
    return argparse.Namespace(**d)


def batched_summary(pkg, args):
    """The batched part of the metric for the default line: 4,096 uneven tries
    in one multi-trie pass against one plain call per trie."""
    r = bench_batched(pkg, _sub_args(args, steps=5, warmup=2, cpu_seconds=5.0))
    return {"workload": r["config"]["workload"], "tries": r["config"]["tries"], "keys": r["config"]["keys"],
            "tries_per_s": r["value"], "ms_per_pass": r["ms_per_step"], "mputs_per_s": r["mputs_per_s"],
            "plain_tries_per_s": r["plain"]["tries_per_s"], "plain_ms_per_pass": r["plain"]["ms_per_pass"],
            "graph_tries_per_s": r["graph"]["tries_per_s"], "graph_ms_per_pass": r["graph"]["ms_per_pass"],
            "speedup_vs_plain": r["speedup_vs_plain"], "roots_equal_plain": True,
            "roots_checked_vs_oracle": r["roots_checked_vs_oracle"], "per_step": r["per_step"],
            "cpu_baseline": r["cpu_baseline"]}


def served_summary(pkg, args):
    """The served part of the metric for the default line: continuous
    batching over 1,024 resident tries at 0.7 x its measured saturation, next
    to one plain insert + hash call per request at 0.7 x theirs."""
    r = bench_served(pkg, _sub_args(args, duration=args.served_duration, plain_duration=args.served_plain_duration,
                                    rates=[0.7]))
    run = r["runs"][0]
    return {"workload": r["config"]["workload"], "lanes": r["config"]["lanes"],
            "puts_per_s": run["puts_per_s"], "requests_per_s": run["requests_per_s"],
            "p50_ms": run["p50_ms"], "p99_ms": run["p99_ms"], "requests": run["requests"],
            "requests_per_batch": run["requests_per_batch"], "deferred_by_memory": run["deferred_by_memory"],
            "saturation": r["saturation"],
            "plain": {k: r["plain"][k] for k in ("saturation_requests_per_s", "offered_rps", "requests_per_s",
                                                  "p50_ms", "p99_ms", "how")},
            "speedup_vs_plain_saturation": r["speedup_vs_plain_saturation"],
            "duration_s": args.served_duration, "plain_duration_s": args.served_plain_duration}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="phase", choices=["phase", "batched", "lock", "served", "hbm-copy"])
    ap.add_argument("--duration", type=float, default=60.0, help="served: seconds per offered rate")
    ap.add_argument("--plain-duration", type=float, default=20.0, help="served: seconds of the plain run")
    ap.add_argument("--rates", type=float, nargs="+", default=[0.7], help="served: fractions of saturation")
    ap.add_argument("--lanes", type=int, default=4, help="served: concurrent lanes (forests)")
    ap.add_argument("--max-batch-keys", type=int, default=1 << 20, help="served: keys per lane batch")
    ap.add_argument("--pool", type=int, default=1500, help="served: distinct pre-generated requests")
    ap.add_argument("--backlog", type=int, default=400, help="served: requests of the saturation run")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="batched: seconds of the CPU baseline sample")
    ap.add_argument("--no-batched", action="store_true", help="phase: leave out the batched sub-object")
    ap.add_argument("--no-served", action="store_true", help="phase: leave out the served sub-object")
    ap.add_argument("--served-duration", type=float, default=10.0,
                    help="phase: seconds of the served sub-object's offered-load run")
    ap.add_argument("--served-plain-duration", type=float, default=5.0,
                    help="phase: seconds of the served sub-object's plain-call run")
    ap.add_argument("--spread-max", type=int, default=None,
                    help="phase: fixed hc_spread_max instead of the autotuner's choice")
    ap.add_argument("--pool-host", action="store_true",
                    help="phase: node pools, key and value arenas in pinned host memory (tries beyond HBM)")
    ap.add_argument("--no-fuse-inner", action="store_true",
                    help="phase: inner levels as 16-lane encode kernel + HBM scratch + permutation kernel")
    ap.add_argument("--no-fuse-leaves", action="store_true",
                    help="phase: leaf level as encode kernel + HBM scratch + permutation kernel")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        raise SystemExit(2)
    pkg = _import_package()
    out = {"phase": bench_phase, "batched": bench_batched, "lock": bench_lock,
           "served": bench_served, "hbm-copy": bench_hbm_copy}[args.config](pkg, args)
    if args.config == "phase":
        # the metric's other two parts: batched vs plain tries/s, served p99
        if not args.no_batched:
            print("bench: batched sub-object", file=sys.stderr, flush=True)
            out["batched"] = batched_summary(pkg, args)
        if not args.no_served:
            print("bench: served sub-object", file=sys.stderr, flush=True)
            out["served"] = served_summary(pkg, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
