"""The oracle pinned against the reference's golden vectors, the reference's own CPU code
(oracle/_ref, where built), brute force and closed forms -- and shown to catch planted mistakes."""
import os

import numpy as np
import pytest

import oracle
from harmonia_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U64 = np.uint64
KEY_MAX = U64(0xFFFFFFFFFFFFFFFF)


def load_gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def layout_for(keys, vals, F=64, fill=8):
    order = np.argsort(keys, kind="stable")
    return oracle.build_bfs(keys[order], vals[order], F, fill)


# ---------------------------------------------------------------- golden vectors (reference outputs)
@pytest.mark.parametrize("name,F", [("search_rbpt64.npz", 64), ("search_rbpt128.npz", 128)])
@pytest.mark.parametrize("fill", [8, 5, 1])
def test_search_matches_reference_golden(name, F, fill):
    g = load_gold(name)
    L = layout_for(g["keys"], g["vals"], F, fill)
    assert oracle.check_layout(L) == 0
    v, f = oracle.search_cfb(L, g["queries"])
    assert np.array_equal(f, g["rbpt_found"])
    assert np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])
    assert np.all(v[f == 0] == 0)
    if "harmonia_vals" in g:
        h = g["harmonia_vals"]
        assert np.array_equal(np.where(f == 1, v.view(np.int64), -1), h)


@pytest.mark.parametrize("name,F", [("search_rbpt64.npz", 64), ("search_rbpt128.npz", 128)])
@pytest.mark.parametrize("fill", [8, 5, 1])
@pytest.mark.parametrize("avx2", [True, False])
def test_search_harmonia_matches_reference_golden(name, F, fill, avx2):
    g = load_gold(name)
    L = layout_for(g["keys"], g["vals"], F, fill)
    v, f = oracle.search_harmonia(L, g["queries"], 2, avx2=avx2)
    assert np.array_equal(f, g["rbpt_found"])
    assert np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])
    assert np.all(v[f == 0] == 0)
    if "harmonia_vals" in g:
        assert np.array_equal(np.where(f == 1, v.view(np.int64), -1), g["harmonia_vals"])


def test_search_harmonia_edge_keys():
    """Queries equal to separators and index entries, one above and below them, and the key-space
    ends: the AVX2 signed compares on top-bit-flipped keys must order like unsigned keys."""
    keys = np.concatenate([np.array([0, 1, (1 << 63) - 1, 1 << 63, (1 << 64) - 2], dtype=U64),
                           W.gen_keys(5000, 91)[0]])
    keys = np.unique(keys)
    sk, sv = keys, W.splitmix64_mix(keys)
    for F in (32, 64, 128):
        L = oracle.build_bfs(sk, sv, F, 3)
        seps = L.inner[L.inner != KEY_MAX]
        q = np.concatenate([seps, seps - U64(1), seps + U64(1), sk[:50], np.array([0, KEY_MAX, KEY_MAX - U64(1)],
                                                                                 dtype=U64)])
        ev, ef = oracle.search_sorted(sk, sv, q)
        for avx2 in (True, False):
            v, f = oracle.search_harmonia(L, q, 2, avx2=avx2)
            assert np.array_equal(f, ef) and np.array_equal(v, ev)


def test_search_sorted_matches_reference_golden():
    g = load_gold("search_rbpt64.npz")
    sk = np.sort(g["keys"])
    v, f = oracle.search_sorted(sk, W.splitmix64_mix(sk), g["queries"])
    assert np.array_equal(f, g["rbpt_found"]) and np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])


def test_range_matches_reference_golden():
    g = load_gold("range_rbpt64.npz")
    sk = np.sort(g["keys"])
    sv = W.splitmix64_mix(sk)
    # regular_bpt find_range drops ranges whose start lies above every key of its leaf
    # (H/regular_bpt/regular_bpt.cpp:325-326): the two reference trees differ only there
    rc, hc = g["rbpt_counts"], g["harmonia_counts"]
    diff = np.nonzero(rc != hc)[0]
    assert np.all(rc[diff] == 0) and len(diff) < len(rc) // 20
    cap = int(hc.max()) + 1
    for fill in (8, 3):
        L = oracle.build_bfs(sk, sv, 64, fill)
        _, k, v, c = oracle.range_walk(L, g["lo"], his=g["hi"], cap=cap)
        assert np.array_equal(c, hc.astype(np.uint32))
        got_k = np.concatenate([k[i * cap:i * cap + c[i]] for i in range(len(c))])
        got_v = np.concatenate([v[i * cap:i * cap + c[i]] for i in range(len(c))])
        assert np.array_equal(got_k, g["harmonia_keys"]) and np.array_equal(got_v, g["harmonia_vals"])
        same = rc == hc  # where the regular tree answers, it agrees too
        got_r = np.concatenate([k[i * cap:i * cap + c[i]] for i in range(len(c)) if same[i]])
        exp_r = np.concatenate([g["rbpt_keys"][int(rc[:i].sum()):int(rc[:i].sum()) + rc[i]]
                                for i in range(len(rc)) if same[i]])
        assert np.array_equal(got_r, exp_r)
        off, k, v, c = oracle.range_walk(L, g["starts"], lens=g["lens"])
        assert np.array_equal(c, g["len_counts"])
        got_k = np.concatenate([k[int(off[i]):int(off[i]) + c[i]] for i in range(len(c))])
        assert np.array_equal(got_k, g["len_keys"])


def test_update_closed_form_matches_reference_golden():
    g = load_gold("update_rbpt64.npz")
    bk, bv = g["base_keys"], g["base_vals"]
    o = np.argsort(bk)
    k2, v2 = oracle.upsert_sorted(bk[o], bv[o], g["new_keys"], g["new_vals"])
    v, f = oracle.search_sorted(k2, v2, g["queries"])
    assert np.array_equal(f, g["rbpt_found"]) and np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])
    L = oracle.build_bfs(k2, v2, 64, 6)
    v, f = oracle.search_cfb(L, g["queries"])
    assert np.array_equal(f, g["rbpt_found"]) and np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])


# ---------------------------------------------------------------- live reference (oracle/_ref)
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("order,F", [(64, 64), (128, 128), (64, 32)])
def test_search_matches_reference_code(order, F):
    keys, vals = W.gen_keys(30000, 17 + order)
    R = oracle.RefRegularTree(keys, vals, order)
    sk = np.sort(keys)
    q = W.point_queries(sk, 30000, 3)
    rv, rf = R.find(q)
    L = oracle.build_bfs(sk, W.splitmix64_mix(sk), F, 7)
    v, f = oracle.search_cfb(L, q)
    assert np.array_equal(f, rf) and np.array_equal(v[f == 1], rv[rf == 1])
    H = oracle.RefHarmoniaTree(keys, vals)
    assert np.array_equal(np.where(f == 1, v.view(np.int64), -1), H.search(q))
    hv, hf = oracle.search_harmonia(L, q, 2)
    assert np.array_equal(hf, f) and np.array_equal(hv, v)


@needs_ref
def test_cpu_config_2p20_reference_paths_agree():
    """The CPU configuration (2^20 keys, fanout 64, 2^20 queries: 90% hits, 10% misses, unsorted):
    the reference's own regular_bpt find and Harmonia BPlusTree::search under their OpenMP loops
    (search_cpu_multithread, search_cpu_test), the two restatements over the bulk-built layout and
    the closed form give the same value / not-found per query."""
    seed = W.config_seed(0)
    keys, vals = W.gen_keys(1 << 20, seed)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    q = W.point_queries(sk, 1 << 20, seed + 1)
    ev, ef = oracle.search_sorted(sk, sv, q)
    assert int(ef.sum()) == 943718
    L = oracle.build_bfs(sk, sv, 64, 8)
    for fn in (oracle.search_cfb, oracle.search_harmonia):
        v, f = fn(L, q, 4)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
    rv, rf = oracle.RefRegularTree(keys, vals, 64).find(q, nthreads=4)
    assert np.array_equal(rf, ef) and np.array_equal(rv[rf == 1], ev[ef == 1])
    hv = oracle.RefHarmoniaTree(keys, vals).search(q, nthreads=4)
    assert np.array_equal(np.where(ef == 1, ev.view(np.int64), -1), hv)


@needs_ref
def test_range_matches_reference_code():
    keys, vals = W.gen_keys(20000, 31)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    R = oracle.RefRegularTree(keys, vals, 64)
    H = oracle.RefHarmoniaTree(keys, vals)
    L = oracle.build_bfs(sk, sv, 64, 8)
    g = W.rng(5, 5)
    for _ in range(40):
        lo = U64(g.integers(0, 1 << 63, dtype=np.uint64))
        hi = lo + U64(g.integers(0, 1 << 53, dtype=np.uint64))
        rk, rv = R.find_range(lo, hi, 20000)
        hk, hv = H.range_search(lo, hi, 20000)
        _, k, v, c = oracle.range_walk(L, [lo], his=[hi], cap=20000)
        assert c[0] == len(hk) and np.array_equal(k[:c[0]], hk) and np.array_equal(v[:c[0]], hv)
        if len(rk):  # regular find_range may return nothing (leaf-boundary quirk), else it agrees
            assert np.array_equal(k[:c[0]], rk) and np.array_equal(v[:c[0]], rv)


@needs_ref
def test_inserts_match_reference_code():
    keys, vals = W.gen_keys(6000, 41)
    R = oracle.RefRegularTree(keys[:5000], vals[:5000], 64)
    R.insert(keys[5000:], vals[5000:] ^ U64(7))
    o = np.argsort(keys[:5000])
    k2, v2 = oracle.upsert_sorted(keys[:5000][o], vals[:5000][o], keys[5000:], vals[5000:] ^ U64(7))
    q = np.concatenate([keys, W.gen_misses(6000, 500, 41)])
    rv, rf = R.find(q)
    v, f = oracle.search_sorted(k2, v2, q)
    assert np.array_equal(f, rf) and np.array_equal(v, np.where(rf == 1, rv, 0))


# ---------------------------------------------------------------- brute force / closed forms
@pytest.mark.parametrize("n", [0, 1, 2, 7, 8, 9, 63, 64, 65, 511, 513, 4097])
@pytest.mark.parametrize("F,fill", [(32, 8), (64, 8), (128, 1), (64, 3)])
def test_search_brute_force(n, F, fill):
    keys, vals = W.gen_keys(n, n * 7 + F)
    d = dict(zip(keys.tolist(), vals.tolist()))
    sk = np.sort(keys)
    L = oracle.build_bfs(sk, W.splitmix64_mix(sk), F, fill)
    assert oracle.check_layout(L) == 0
    extra = [0, 1, (1 << 63) - 2, (1 << 64) - 2, int(KEY_MAX)]
    if n:
        extra += [int(sk[0]) - 1, int(sk[-1]) + 1]
    q = np.concatenate([W.point_queries(sk, 300, n), np.array(extra, dtype=U64)])
    v, f = oracle.search_cfb(L, q)
    for qi, vi, fi in zip(q.tolist(), v.tolist(), f.tolist()):
        assert fi == (qi in d) and (not fi or vi == d[qi])


def test_bfs_layout_brute_force_small():
    """Tiny tree laid out by hand: 20 keys, F=32, fill 2 -> 10 leaves under one root."""
    sk = np.arange(10, 210, 10, dtype=U64)
    L = oracle.build_bfs(sk, sk + U64(1), 32, 2)
    assert L.height == 1 and L.n_inner == 1 and L.n_leaves == 10
    assert L.prefix.tolist() == [1]
    node = L.inner[0]
    seps = node[8:8 + 32]
    assert seps[:9].tolist() == [30, 50, 70, 90, 110, 130, 150, 170, 190]
    assert np.all(seps[9:] == KEY_MAX)
    assert node[:8].tolist()[:4] == [int(seps[7]), int(seps[15]), int(seps[23]), int(seps[31])]
    assert L.leaves[0, :2].tolist() == [10, 20] and np.all(L.leaves[0, 2:8] == KEY_MAX)
    assert L.leaves[0, 8:10].tolist() == [11, 21]


def test_prefix_recurrence_two_levels():
    sk = np.arange(1, 1 + 8 * 40, dtype=U64)  # 40 full leaves, F=32 -> 2 last-level nodes + root
    L = oracle.build_bfs(sk, sk, 32, 8)
    assert L.height == 2 and L.n_inner == 3
    assert L.prefix.tolist() == [1, 3, 35]  # root -> nodes 1,2; node 1 -> leaves 0..31 (BFS 3..34); node 2 -> 32..
    assert L.inner[0, 8] == sk[32 * 8]  # root separator = first key of the second subtree


def test_psa_sort_brute_force():
    g = W.rng(1, 1)
    for n in (0, 1, 5, 100, 3000):
        keys = g.integers(0, 1 << 63, n, dtype=np.uint64)
        keys[: n // 3] = keys[n // 2: n // 2 + n // 3] if n > 3 else keys[: n // 3]  # repeated keys
        for b, e in ((48, 64), (56, 64), (0, 64), (60, 64), (32, 40), (10, 10)):
            ko, po = oracle.psa_sort(keys, b, e)
            mask = (1 << (e - b)) - 1
            exp = sorted(range(n), key=lambda i: ((int(keys[i]) >> b) & mask, i))
            assert po.tolist() == exp and np.array_equal(ko, keys[po])


def test_upsert_sorted_dict_semantics():
    g = W.rng(2, 2)
    base = np.unique(g.integers(0, 1000, 300, dtype=np.uint64))
    bv = base * U64(3)
    bk = g.integers(0, 1200, 500, dtype=np.uint64)
    bvals = g.integers(0, 1 << 60, 500, dtype=np.uint64)
    d = dict(zip(base.tolist(), bv.tolist()))
    for k, v in zip(bk.tolist(), bvals.tolist()):
        d[k] = v  # later writes win
    k2, v2 = oracle.upsert_sorted(base, bv, bk, bvals)
    assert k2.tolist() == sorted(d) and v2.tolist() == [d[k] for k in sorted(d)]


def test_range_closed_form_vs_walk():
    keys, vals = W.gen_keys(50000, 9)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    starts, lens = W.range_queries(2000, 9)
    starts[:10] = sk[:10]
    starts[10] = sk[-1] + U64(1)
    for F, fill in ((64, 8), (128, 4), (32, 1)):
        L = oracle.build_bfs(sk, sv, F, fill)
        a = oracle.range_walk(L, starts, lens=lens)
        b = oracle.range_sorted(sk, sv, starts, lens=lens)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


# ---------------------------------------------------------------- planted mistakes are caught
def test_planted_child_rule_mistake_fails_golden():
    """A lower_bound child rule (k > target instead of k >= target) must disagree with the
    reference golden vectors: separators equal to a query then route to the wrong leaf."""
    g = load_gold("search_rbpt64.npz")
    sk = np.sort(g["keys"])
    L = oracle.build_bfs(sk, W.splitmix64_mix(sk), 64, 8)
    # emulate the mistake: shift every separator up by one, so key == separator goes left
    bad = oracle.Layout(L.inner.copy(), L.prefix, L.leaves, L.height, L.fanout, L.level_start)
    real = bad.inner != KEY_MAX
    bad.inner[real] += U64(1)
    v, f = oracle.search_cfb(bad, g["queries"])
    assert not np.array_equal(f, g["rbpt_found"])  # the golden vectors expose the mistake
    v, f = oracle.search_cfb(L, g["queries"])
    assert np.array_equal(f, g["rbpt_found"])  # the unmutated layout does not
    assert oracle.check_layout(bad) != 0


def test_planted_index_step_mistake_caught():
    """search_harmonia's index step compares index slots 4..7 only when slots 0..3 are all <= key
    (getChildIdx_avx2_version2).  Dropping that second half (emulated by padding slots 4..7, so
    the step never counts them) sends every query of the upper half of a node to group 4 at most;
    the reference golden vectors must catch it, while the key-only scan (search_cfb) is unaffected."""
    g = load_gold("search_rbpt64.npz")
    sk = np.sort(g["keys"])
    L = oracle.build_bfs(sk, W.splitmix64_mix(sk), 64, 8)
    bad = oracle.Layout(L.inner.copy(), L.prefix, L.leaves, L.height, L.fanout, L.level_start)
    bad.inner.reshape(-1, oracle.node_slots(64))[:, 4:8] = KEY_MAX
    v, f = oracle.search_harmonia(bad, g["queries"], 2)
    assert not np.array_equal(f, g["rbpt_found"])
    v, f = oracle.search_cfb(bad, g["queries"])
    assert np.array_equal(f, g["rbpt_found"]) and np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])


def test_planted_prefix_mistake_caught():
    keys, vals = W.gen_keys(20000, 77)
    sk = np.sort(keys)
    L = oracle.build_bfs(sk, W.splitmix64_mix(sk), 64, 8)
    bad = oracle.Layout(L.inner, L.prefix.copy(), L.leaves, L.height, L.fanout, L.level_start)
    bad.prefix[2] += 1  # off-by-one in one prefix entry
    assert oracle.check_layout(bad) != 0
    q = W.point_queries(sk, 20000, 1)
    v, f = oracle.search_cfb(bad, q)
    ev, ef = oracle.search_sorted(sk, W.splitmix64_mix(sk), q)
    assert not np.array_equal(f, ef)


def test_planted_unstable_sort_caught():
    keys = np.array([5 << 48, 3 << 48, (5 << 48) + 1, 3 << 48, 5 << 48], dtype=U64)
    ko, po = oracle.psa_sort(keys, 48, 64)
    assert po.tolist() == [1, 3, 0, 2, 4]  # equal digits keep input order
    assert po.tolist() != [3, 1, 0, 2, 4]


def test_value_rule_pinned():
    # splitmix64 published test vector: seed 1234567 -> first output 6457827717110365317
    assert int(W.splitmix64_stream(1234567, 1)[0]) == 6457827717110365317
    assert oracle.splitmix64_mix(1234567) == 6457827717110365317
    k = np.array([0, 1, 1 << 62], dtype=U64)
    assert [oracle.splitmix64_mix(int(x)) for x in k] == W.splitmix64_mix(k).tolist()
