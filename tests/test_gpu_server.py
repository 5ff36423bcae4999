"""The serving loop: each request's result is the same alone and in a mix of concurrent requests,
reads see a whole epoch (before or after an update, never between), and admission honours the
device-memory budget."""
import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree, HarmoniaError, Server
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
U64 = np.uint64


@pytest.fixture(scope="module")
def tree_and_keys(gpu):
    keys, vals = W.gen_keys(400000, 21)
    t = BPlusTree.from_keys(keys, vals, leaf_fill=6)
    sk = np.sort(keys)
    return t, sk, W.splitmix64_mix(sk)


def make_requests(sk, seed, count):
    g = W.rng(seed, 1)
    reqs = []
    for i in range(count):
        if g.random() < 0.85:
            m = int(W.log_uniform_sizes(1, 1, 20000, seed + i)[0])
            reqs.append(("lookup", W.point_queries(sk, m, seed + 7 * i)))
        else:
            starts, lens = W.range_queries(int(g.integers(1, 8)), seed + i)
            reqs.append(("range", (starts, lens)))
    return reqs


def check_lookup(sk, sv, q, out):
    v, f = out
    ev, ef = oracle.search_sorted(sk, sv, q)
    return np.array_equal(f, ef) and np.array_equal(v, ev)


def check_range(sk, sv, starts, lens, out):
    off, k, v, c = out
    eo, ek, ev, ec = oracle.range_sorted(sk, sv, starts, lens=lens)
    if not np.array_equal(c, ec):
        return False
    for i in range(len(c)):
        a, b = int(off[i]), int(off[i]) + int(c[i])
        if not (np.array_equal(k[a:b], ek[a:b]) and np.array_equal(v[a:b], ev[a:b])):
            return False
    return True


@pytest.mark.parametrize("coalesce", [True, False])
def test_alone_equals_mixed(tree_and_keys, coalesce):
    t, sk, sv = tree_and_keys
    reqs = make_requests(sk, 5, 60)
    srv = Server(t, max_inflight=4, coalesce=coalesce)
    alone = []
    for kind, payload in reqs:  # one at a time
        tk = srv.lookup(payload) if kind == "lookup" else srv.range(*payload)
        st, ep, out = srv.wait(tk)
        assert st == 0
        alone.append(out)
    tickets = []
    for kind, payload in reqs:  # all at once, interleaved with extra traffic
        tickets.append(srv.lookup(payload) if kind == "lookup" else srv.range(*payload))
        srv.lookup(W.point_queries(sk, 3000, len(tickets)))
    for (kind, payload), tk, a in zip(reqs, tickets, alone):
        st, ep, out = srv.wait(tk)
        assert st == 0
        for x, y in zip(a, out):
            assert np.array_equal(x, y)
        assert check_lookup(sk, sv, payload, out) if kind == "lookup" else check_range(sk, sv, *payload, out)
    srv.drain()
    s = srv.stats()
    assert s["completed"] >= 3 * len(reqs) - 5 and s["p99_us"] >= s["p50_us"] > 0
    srv.close()


def test_coalescing_under_load(tree_and_keys):
    """With one worker busy, requests that arrive meanwhile join the next batch: fewer launches than
    requests, and every result is still exact."""
    t, sk, sv = tree_and_keys
    srv = Server(t, max_inflight=1)
    qs = [W.point_queries(sk, 5000 + 37 * i, 300 + i) for i in range(200)]
    tickets = [srv.lookup(q) for q in qs]
    for q, tk in zip(qs, tickets):
        st, ep, out = srv.wait(tk)
        assert st == 0 and check_lookup(sk, sv, q, out)
    s = srv.stats()
    assert s["completed"] == 200 and s["launches"] < 200
    srv.close()


def test_updates_between_epochs(gpu):
    """Reads interleaved with update batches: every read matches the whole state of the epoch it
    reports (initial keys plus, in submission order, every update published up to that epoch), and
    a read submitted after an update completed sees it."""
    keys, vals = W.gen_keys(100000, 31)
    t = BPlusTree.from_keys(keys, vals, leaf_fill=6)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    srv = Server(t, max_inflight=3)
    new = W.gen_misses(100000, 20000, 31)
    nv = W.splitmix64_mix(new ^ U64(5))
    probe = np.concatenate([sk[:2000], new])
    tickets = []
    for step in range(5):
        part = slice(step * 4000, (step + 1) * 4000)
        tickets.append(("read", srv.lookup(probe)))
        tickets.append(("update", srv.update(new[part], nv[part]), part))
    tickets.append(("read", srv.lookup(probe)))
    reads, ups = [], []
    for item in tickets:
        st, ep, out = srv.wait(item[1])
        assert st == 0
        (reads if item[0] == "read" else ups).append((ep, out if item[0] == "read" else item[2]))
    states = {0: (sk, sv)}
    cur = (sk, sv)
    for ep in sorted({e for e, _ in ups}):
        for e, part in ups:  # submission order within the epoch
            if e == ep:
                cur = oracle.upsert_sorted(cur[0], cur[1], new[part], nv[part])
        states[ep] = cur
    for ep, out in reads:
        k2, v2 = states[max(e for e in states if e <= ep)]
        assert check_lookup(k2, v2, probe, out), ep
    # after every update completed, a new read sees all of them
    final = oracle.upsert_sorted(sk, sv, new, nv)
    st, ep, out = srv.wait(srv.lookup(probe))
    assert st == 0 and check_lookup(final[0], final[1], probe, out)
    srv.drain()
    assert srv.stats()["epochs"] >= 1
    srv.close()


def test_budget_rejects_oversized_request(tree_and_keys):
    t, sk, sv = tree_and_keys
    srv = Server(t, mem_budget=1 << 20)
    with pytest.raises(HarmoniaError):
        srv.lookup(W.point_queries(sk, 200000, 1))
    tk = srv.lookup(W.point_queries(sk, 100, 2))
    assert srv.wait(tk)[0] == 0
    assert srv.stats()["rejected"] == 1
    srv.close()
