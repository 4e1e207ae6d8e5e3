"""Continuous-batching server: every request's post-state root is the root the
trie has when that trie's requests are replayed alone, in order (bit-exact),
whatever else shares its batches; admission by the memory budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(workloads, seed, count, tries, median=150, cap=3000):
    rng = np.random.default_rng(seed)
    targets = workloads.zipf_ranks(rng, tries, 0.8, count)
    sizes = np.clip(np.round(np.exp(np.log(median) + rng.standard_normal(count))), 1, cap).astype(int)
    reqs = []
    seen = {}
    for t, s in zip(targets, sizes):
        keys = workloads.unique_random_keys(rng, int(s), 32)
        prev = seen.get(int(t))
        if prev is not None and len(prev) > 4:  # some updates of the trie's earlier keys
            k = min(len(prev), int(s)) // 3
            keys[:k] = prev[:k]
        flat, lens = workloads.random_values(rng, int(s), 20, 110)
        reqs.append((int(t), workloads.batch_from_fixed(keys, flat, lens)))
        seen[int(t)] = keys
    return reqs


def _alone(gmpt, reqs, mode):
    """Post-state root of every request, each trie replayed alone."""
    roots = [None] * len(reqs)
    by_trie = {}
    for i, (t, b) in enumerate(reqs):
        by_trie.setdefault(t, []).append(i)
    m = gmpt.MPT(device=0)
    for t, idx in by_trie.items():
        m.clear()
        for i in idx:
            put = m.puts_2phase_with_valuehp if mode == 0 else m.puts_latching_with_valuehp_v2
            targets, n = put(reqs[i][1])
            m.hash_onepass_v2(targets, n)
            roots[i] = m.get_root_hash()
    m.close()
    return roots


@pytest.mark.parametrize("mode", [0, 1], ids=["phasenu", "locknu"])
def test_results_same_alone_and_in_mix(gmpt, workloads, oracle, mode):
    tries = 24
    reqs = _requests(workloads, 11 + mode, 160, tries)
    srv = gmpt.Server(device=0, num_tries=tries, lanes=3, mode=mode)
    ids = [srv.submit(t, b) for t, b in reqs]
    got = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] < len(reqs)  # requests really were coalesced
    assert got == _alone(gmpt, reqs, mode)
    # final state of two tries against the oracle
    for t in sorted({t for t, _ in reqs})[:2]:
        o = oracle.OracleMPT()
        last = None
        for i, (tt, b) in enumerate(reqs):
            if tt == t:
                o.puts(b)
                last = i
        o.hash()
        assert got[last] == o.root()


def test_memory_budget_admission(gmpt, workloads):
    """Requests run while the lane's pool bytes fit the budget; once the tries
    have filled it, the next request fails cleanly and earlier results stand."""
    reqs = _requests(workloads, 21, 80, 8, median=400)
    srv = gmpt.Server(device=0, num_tries=8, lanes=1, mem_budget=8 << 20)
    big = workloads.config1_oracle(seed=3, n=100_000)  # its estimate alone exceeds the budget
    with pytest.raises(gmpt.GmptError):
        srv.submit(0, big)
    roots, failed_at = [], None
    for i, (t, b) in enumerate(reqs):
        rid = srv.submit(t, b)
        try:
            roots.append(srv.wait(rid)[0])
        except gmpt.GmptError as e:
            assert "budget" in str(e)
            failed_at = i
            break
    st = srv.stats()
    srv.close()
    assert failed_at is not None and failed_at > 0 and st["failed"] == 1
    assert roots == _alone(gmpt, reqs[:failed_at], 0)


def test_malformed_request_refused_others_unaffected(gmpt, workloads):
    reqs = _requests(workloads, 31, 20, 4)
    srv = gmpt.Server(device=0, num_tries=4, lanes=2)
    bad = workloads.batch_from_hex([bytes([1, 2, 16]), bytes([1, 17, 16])], [b"x" * 40, b"y" * 40])
    with pytest.raises(gmpt.GmptError):
        srv.submit(1, bad)
    ids = [srv.submit(t, b) for t, b in reqs]
    got = [srv.wait(i)[0] for i in ids]
    srv.close()
    assert got == _alone(gmpt, reqs, 0)


def test_latching_server_updates_stay_within_budget(gmpt, workloads, oracle):
    """A LockNU server lane rewriting the same keys over and over: each
    update keeps only its staged value bytes, so thirty rounds of updates to
    a 3,000-key trie stay within an 8 MB budget (a node and a packed key per
    update would need about 30 MB), and the final root is the oracle's."""
    rng = np.random.default_rng(41)
    keys = workloads.unique_random_keys(rng, 3000, 32)
    srv = gmpt.Server(device=0, num_tries=1, lanes=1, mem_budget=8 << 20, mode=1)
    o = oracle.OracleMPT()
    for r in range(31):
        flat, lens = workloads.random_values(rng, len(keys), 20, 40)
        b = workloads.batch_from_fixed(keys, flat, lens)
        root, _ = srv.wait(srv.submit(0, b))
        o.puts(b)
    o.hash()
    st = srv.stats()
    srv.close()
    assert st["completed"] == 31 and st["failed"] == 0
    assert root == o.root()
