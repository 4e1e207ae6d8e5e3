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

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "gpu-merkle-patricia-trie_amd")

# MI355X published peaks used for the roofline
HBM_PEAK_GBS = 8000.0
CUS, LANES, PEAK_CLOCK_GHZ = 256, 64, 2.4
# 64-bit logical ops per second: the VALU is 32-bit, a 64-bit xor/and/rotate
# takes two lane-ops
VALU64_PEAK_TOPS = CUS * LANES * PEAK_CLOCK_GHZ * 1e9 / 2 / 1e12
OPS64_PER_PERM = 3720  # Keccak-f[1600]: 155 64-bit lane-ops per round x 24 rounds


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


def cpu_baseline_build(batch, label):
    """The oracle (C restatement of the reference CPU MPT) timed on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    secs, root = oracle.time_build(batch)
    return {"value": batch.n / secs, "unit": "puts+root-hash/s", "cores": 1, "kind": "port",
            "sample": f"{label}: full workload, puts_baseline + hashs_dirty_flag, 1 run, {secs:.2f} s",
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}, root


def bench_phase(pkg, args):
    wl = pkg.workloads
    t0 = time.time()
    b = wl.config2_phase(n=args.n)
    gen_s = time.time() - t0
    db = pkg.DeviceBatch(b, adopt_values=True)
    t = pkg.MPT(device=0)
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
    hash_ms = []
    t.timer_start()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        hash_ms.append(t.stats()["last_hash_ms"])
    t.synchronize()
    wall = time.perf_counter() - w0
    dev_ms = t.timer_stop()
    st = t.stats()
    ms_per_step = wall * 1e3 / args.steps
    value = b.n * args.steps / wall

    # roofline of the dominant kernel (PhaseHC's fused RLP encode + Keccak)
    perms = st["perms"]
    hms = sum(hash_ms) / len(hash_ms)
    achieved = OPS64_PER_PERM * perms / (hms * 1e-3) / 1e12
    # whole-path roofline: input bytes + node bytes over HBM, Keccak over VALU
    b_in = b.key_bytes + 8 * b.n + b.value_bytes + 16 * b.n + 8 * b.n
    b_nodes = st["enc_bytes"] + 32 * st["hashed_nodes"]
    t_roof = (b_in + b_nodes) / (HBM_PEAK_GBS * 1e9) + OPS64_PER_PERM * perms / (VALU64_PEAK_TOPS * 1e12)

    out = {
        "metric": "MPT puts+root-hash ops/s at 1M 32B keys", "value": value, "unit": "puts+root-hash/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8/u64 (integer)",
        "data": "synthetic (seeded uniform 32-byte keys, 70-110-byte random values)",
        "config": {"workload": "PhaseNU+PhaseHC, 1,000,000 unique 32-byte keys, 70-110-byte values, "
                               "one batch into an empty trie, inputs resident in HBM",
                   "keys": b.n, "value_bytes": b.value_bytes},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU64_PEAK_TOPS, "unit": "T64op/s",
                     "frac": achieved / VALU64_PEAK_TOPS, "traffic": None,
                     "kernel": "k_hc_hash (RLP encode + Keccak-f[1600], one state per lane)",
                     "per_step": {"permutations": perms, "nodes_hashed": st["hashed_nodes"],
                                  "hash_ms": hms, "launches": st["hash_launches"]}},
        "path_roofline": {"t_roof_ms": t_roof * 1e3, "t_measured_ms": ms_per_step,
                          "frac": t_roof * 1e3 / ms_per_step, "bytes_in": b_in, "bytes_nodes": b_nodes,
                          "device_ms_per_step": dev_ms / args.steps},
        "mputs_per_s": value / 1e6, "mhashes_per_s": st["hashed_nodes"] / (ms_per_step * 1e-3) / 1e6,
        "root": roots[-1].hex(),
    }
    if len(set(roots)) != 1:
        raise SystemExit(f"bench: roots differ between steps: {set(r.hex() for r in roots)}")
    if not args.no_cpu_baseline:
        cb, croot = cpu_baseline_build(b, "1,000,000 keys (the same batch)")
        if croot != roots[-1]:
            raise SystemExit(f"bench: GPU root {roots[-1].hex()} != oracle root {croot.hex()}")
        out["cpu_baseline"] = cb
        out["root_matches_oracle"] = True
    else:
        out["cpu_baseline"] = None
    out["gen_s"] = gen_s
    db.free()
    t.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="phase", choices=["phase"])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        raise SystemExit(2)
    pkg = _import_package()
    out = bench_phase(pkg, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
