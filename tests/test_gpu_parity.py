"""GPU parity: every result of the HIP path is compared bit for bit with the oracle, the reference's
golden vectors and closed forms.  All calls go through the C ABI (libharmonia_amd)."""
import os

import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree, HarmoniaError, psa_sort
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U64 = np.uint64
KEY_MAX = U64(0xFFFFFFFFFFFFFFFF)


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def make(n, F=64, fill=8, seed=1, **kw):
    keys, vals = W.gen_keys(n, seed)
    t = BPlusTree.from_keys(keys, vals, fanout=F, leaf_fill=fill, **kw)
    sk = np.sort(keys)
    return t, keys, vals, sk, W.splitmix64_mix(sk)


def edge(sk):
    e = [0, 1, int(KEY_MAX), int(KEY_MAX) - 1, (1 << 63) - 1, (1 << 63) - 2]
    if len(sk):
        e += [int(sk[0]), int(sk[-1]), int(sk[0]) + 1, int(sk[-1]) + 1, max(int(sk[0]) - 1, 0)]
    return np.array(e, dtype=U64)


def queries(sk, m, seed):
    base = W.point_queries(sk, m, seed) if len(sk) else W.gen_misses(0, m, seed)
    return np.concatenate([base, edge(sk)])


# ---------------------------------------------------------------- layout
@pytest.mark.parametrize("n,F,fill", [(0, 64, 8), (1, 64, 8), (8, 32, 8), (9, 32, 8), (513, 64, 8), (4096, 128, 8),
                                      (100003, 64, 8), (100003, 32, 5), (250000, 128, 6), (70000, 64, 1)])
def test_device_layout_equals_oracle(gpu, n, F, fill):
    print(f"layout n={n} F={F} fill={fill}", flush=True)
    t, keys, vals, sk, sv = make(n, F, fill, seed=n + F)
    e = t.export()
    L = oracle.build_bfs(sk, sv, F, fill)
    s = e["stats"]
    assert s.height == L.height and s.n_inner == L.n_inner and s.n_leaves == L.n_leaves and s.n_keys == n
    assert np.array_equal(e["prefix"], L.prefix)
    assert np.array_equal(e["inner"], L.inner)
    assert np.array_equal(e["leaves"], L.leaves)
    assert oracle.check_layout(oracle.Layout(e["inner"], e["prefix"], e["leaves"], s.height, F,
                                             np.array(s.level_start[:], dtype=np.uint32))) == 0


def test_sorted_input_flag_and_rejections(gpu):
    keys, vals = W.gen_keys(5000, 3)
    sk = np.sort(keys)
    t = BPlusTree.from_keys(sk, W.splitmix64_mix(sk), keys_sorted=True)
    assert np.array_equal(t.export()["leaves"], oracle.build_bfs(sk, W.splitmix64_mix(sk), 64, 8).leaves)
    with pytest.raises(HarmoniaError):  # duplicate keys
        BPlusTree.from_keys(np.concatenate([keys, keys[:1]]), np.concatenate([vals, vals[:1]]))
    with pytest.raises(HarmoniaError):  # reserved padding key
        BPlusTree.from_keys(np.concatenate([keys, [KEY_MAX]]), np.concatenate([vals, [U64(0)]]))
    with pytest.raises(HarmoniaError):  # flagged sorted but not sorted
        BPlusTree.from_keys(keys, vals, keys_sorted=True)


# ---------------------------------------------------------------- point search
@pytest.mark.parametrize("F", [32, 64, 128])
def test_search_all_launch_shapes(gpu, F):
    t, keys, vals, sk, sv = make(100003, F, 8, seed=F)
    q = queries(sk, 50000, F)
    ev, ef = oracle.search_sorted(sk, sv, q)
    L = oracle.build_bfs(sk, sv, F, 8)
    cv, cf = oracle.search_cfb(L, q)
    assert np.array_equal(cv, ev) and np.array_equal(cf, ef)
    for g in (1, 2, 4, 8, 16, 32, 64):
        if g > F:
            continue
        for qpg in (1, 2, 4):
            for lds in (0, 1, 2):
                # psa 2 leaves the batch nearly unsorted: with one lane per query the sorted-input
                # kernel then meets more distinct nodes per wave than it shares and falls back
                for psa in (0, 2, 16):
                    for mode in ((1, 2) if psa else (0,)):
                        for sk_ in ((1, 2) if psa else (0,)):  # one descent per query / leaf window
                            v, f = t.search_batch(q, group_size=g, queries_per_group=qpg, lds_levels=lds,
                                                  psa_bits=psa, output_mode=mode, sorted_kernel=sk_)
                            assert np.array_equal(f, ef), (g, qpg, lds, psa, mode, sk_)
                            assert np.array_equal(v, ev), (g, qpg, lds, psa, mode, sk_)


@pytest.mark.parametrize("F", [32, 64, 128])
def test_search_leaf_window(gpu, F):
    """The leaf-window kernel on dense and sparse batches (chunks within the LDS ring and chunks that
    fall back to one descent per query), skewed and edge-key queries, nearly unsorted input (PSA over
    2 bits: ranges that move backwards between chunks), after in-place inserts and after a re-layout."""
    t, keys, vals, sk, sv = make(300007, F, 6, seed=3 * F)
    cur_k, cur_v = sk, sv
    for step in range(3):
        for m, bits in ((400000, 16), (200000, 24), (60000, 16), (3000, 16), (150000, 2), (100000, 32)):
            base = W.point_queries(cur_k, m, 100 * step + m % 97)
            z = W.zipf_queries(cur_k, m // 3, 7 + step)
            q = np.concatenate([base, z, edge(cur_k)])
            ev, ef = oracle.search_sorted(cur_k, cur_v, q)
            for mode in (1, 3):
                pos = np.empty(len(q), dtype=np.uint32) if mode == 3 else None
                v, f = t.search_batch(q, psa_bits=bits, output_mode=mode, sorted_kernel=2, out_pos=pos)
                if mode == 3:
                    v2, f2 = np.empty_like(v), np.empty_like(f)
                    v2[pos], f2[pos] = v, f
                    v, f = v2, f2
                assert np.array_equal(f, ef), (step, m, bits, mode)
                assert np.array_equal(v, ev), (step, m, bits, mode)
        # step 0 -> 1: in-place inserts (leaf minima below their old minimum are impossible: an insert
        # lands in the leaf whose separator is <= the key); step 1 -> 2: enough inserts to re-lay out
        n_new = 20000 if step == 0 else 400000
        new = W.gen_misses(len(keys), n_new, 50 + step) ^ np.uint64(step << 40)
        new = np.setdiff1d(new, cur_k)
        nv = W.splitmix64_mix(new ^ np.uint64(9))
        t.update(new, nv)
        t.publish()
        cur_k, cur_v = oracle.upsert_sorted(cur_k, cur_v, new, nv)


@pytest.mark.parametrize("F", [32, 64])
def test_search_sorted_span_deep_tree(gpu, F):
    """PSA-sorted batches with two lanes x four queries: each wave shares its upper-level node reads and
    loads the run of last-level nodes its 128 queries reach.  A tall tree runs that over several levels;
    few PSA bits leave waves straddling many nodes (the per-group and direct last-level fallbacks), many
    bits keep them in one node and a short run."""
    t, keys, vals, sk, sv = make(1 << 21 if F == 32 else 1 << 22, F, 8, seed=31 + F)
    q = queries(sk, 1 << 18, 5 + F)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for lds in (0, 1, 2):
        for bits in (3, 8, 16, 30):
            for mode in (1, 2):
                v, f = t.search_batch(q, group_size=2, queries_per_group=4, lds_levels=lds, psa_bits=bits,
                                      output_mode=mode)
                assert np.array_equal(f, ef), (lds, bits, mode)
                assert np.array_equal(v, ev), (lds, bits, mode)


@pytest.mark.parametrize("n", [0, 1, 2, 7, 8, 9, 64, 65, 511, 4097])
def test_search_small_and_ragged(gpu, n):
    t, keys, vals, sk, sv = make(n, 64, 8, seed=n)
    for m in (1, 3, 63, 64, 65, 1000):
        q = queries(sk, m, m)
        ev, ef = oracle.search_sorted(sk, sv, q)
        v, f = t.search_batch(q, psa_bits=16)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
        v, f = t.search_batch(q, psa_bits=0)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
    v, f = t.search_batch(np.empty(0, dtype=U64))
    assert len(v) == 0


@pytest.mark.parametrize("name,F", [("search_rbpt64.npz", 64), ("search_rbpt128.npz", 128)])
@pytest.mark.parametrize("fill", [8, 4])
def test_search_reference_golden(gpu, name, F, fill):
    g = gold(name)
    t = BPlusTree.from_keys(g["keys"], g["vals"], fanout=F, leaf_fill=fill)
    for psa in (0, 16):
        v, f = t.search_batch(g["queries"], psa_bits=psa)
        assert np.array_equal(f, g["rbpt_found"])
        assert np.array_equal(v[f == 1], g["rbpt_vals"][f == 1]) and np.all(v[f == 0] == 0)
        if "harmonia_vals" in g:
            assert np.array_equal(np.where(f == 1, v.view(np.int64), -1), g["harmonia_vals"])


@pytest.mark.parametrize("order", ["unsorted", "runs4096", "sorted"])
def test_search_query_orders(gpu, order):
    t, keys, vals, sk, sv = make(300000, 64, 8, seed=11)
    q = W.point_queries(sk, 200000, 12, order=order)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for psa in (0, 16, 24):
        v, f = t.search_batch(q, psa_bits=psa)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)


def test_search_device_pointers(gpu):
    import torch

    t, keys, vals, sk, sv = make(200000, 64, 8, seed=21)
    q = queries(sk, 100000, 22)
    ev, ef = oracle.search_sorted(sk, sv, q)
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    dv = torch.empty(len(q), dtype=torch.int64, device=gpu)
    df = torch.empty(len(q), dtype=torch.uint8, device=gpu)
    s = torch.cuda.Stream()
    for psa in (0, 16):
        t.search_batch(dq, out_vals=dv, out_found=df, stream=s, psa_bits=psa)
        s.synchronize()
        assert np.array_equal(df.cpu().numpy(), ef)
        assert np.array_equal(dv.cpu().numpy().view(U64), ev)
    # device queries and results at an 8-byte but not 16-byte aligned address (a tensor slice): the
    # PSA histogram then reads the queries with 8-byte loads
    dq1 = torch.from_numpy(np.concatenate([np.zeros(1, U64), q]).view(np.int64)).to(gpu)[1:]
    dv1 = torch.empty(len(q) + 1, dtype=torch.int64, device=gpu)[1:]
    df1 = torch.empty(len(q) + 1, dtype=torch.uint8, device=gpu)[1:]
    assert dq1.data_ptr() % 16 == 8
    for mode in (1, 2):
        t.search_batch(dq1, out_vals=dv1, out_found=df1, stream=s, psa_bits=16, output_mode=mode)
        s.synchronize()
        assert np.array_equal(df1.cpu().numpy(), ef) and np.array_equal(dv1.cpu().numpy().view(U64), ev)


@pytest.mark.parametrize("n_q", [1, 2, 255, 256, 257, 65537, 300001])
def test_output_modes(gpu, n_q):
    """PSA results come back in input order by radix un-permute (1) or direct scatter (2)."""
    t, keys, vals, sk, sv = make(120000, 64, 8, seed=n_q)
    q = queries(sk, n_q, n_q + 1)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for mode in (1, 2):
        for g, qpg in ((1, 1), (1, 2), (2, 4), (4, 2), (8, 4)):
            for bits in (16, 24):
                v, f = t.search_batch(q, psa_bits=bits, output_mode=mode, group_size=g, queries_per_group=qpg)
                assert np.array_equal(f, ef), (mode, g, qpg, bits)
                assert np.array_equal(v, ev), (mode, g, qpg, bits)
    pos = np.empty(len(q), dtype=np.uint32)
    v, f = t.search_batch(q, psa_bits=16, output_mode=3, out_pos=pos)
    assert np.array_equal(np.sort(pos), np.arange(len(q), dtype=np.uint32))
    assert np.array_equal(f, ef[pos]) and np.array_equal(v, ev[pos])


def test_sorted_output_matches_psa_order(gpu):
    t, keys, vals, sk, sv = make(50000, 64, 8, seed=31)
    q = queries(sk, 70000, 32)
    pos = np.empty(len(q), dtype=np.uint32)
    v, f = t.search_batch(q, psa_bits=16, sorted_output=True, out_pos=pos)
    ko, po = oracle.psa_sort(q, 48, 64)
    assert np.array_equal(pos, po)
    ev, ef = oracle.search_sorted(sk, sv, ko)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)


@pytest.mark.parametrize("base_bits", [20, 40, 62])
def test_psa_window_follows_key_range(gpu, base_bits, tmp_path):
    """Input-order results sort the queries by the psa_bits bits below the highest bit in which the tree's
    keys differ.  Keys packed into [2^base_bits, 2^base_bits + 2^30) put that window far below bit 64;
    queries inside, below and above the range, and an update epoch that widens it (keys below and above),
    must all match the oracle, and PSA-order output keeps the reference's window [64 - bits, 64)."""
    g = W.rng(base_bits, 7)
    base = np.uint64(1 << base_bits)
    off = np.unique(g.integers(0, 1 << 30, size=260000, dtype=np.uint64))[:200000]
    keys = base + off
    g.shuffle(keys)
    vals = W.splitmix64_mix(keys)
    t = BPlusTree.from_keys(keys, vals, fanout=64, leaf_fill=6)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)

    def check_range(tree, cur):
        st = tree.stats()
        assert (st.key_min, st.key_max) == (int(cur[0]), int(cur[-1]))
        assert st.psa_window_top == (int(cur[0]) ^ int(cur[-1])).bit_length()

    check_range(t, sk)
    assert t.stats().psa_window_top <= 31

    def mixed_queries(sk, seed):
        h = W.rng(seed, 3)
        lo, hi = int(sk[0]), int(sk[-1])
        parts = [sk[h.integers(0, len(sk), size=60000)],                                   # hits
                 h.integers(lo, hi, size=20000, dtype=np.uint64),                          # in range
                 h.integers(0, lo, size=5000, dtype=np.uint64),                            # below
                 h.integers(hi + 1, int(KEY_MAX), size=5000, dtype=np.uint64),             # above
                 np.array([0, lo, hi, lo - 1, hi + 1, int(KEY_MAX) - 1, int(KEY_MAX)], dtype=np.uint64)]
        q = np.concatenate(parts)
        h.shuffle(q)
        return q

    q = mixed_queries(sk, base_bits)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for mode in (1, 2):
        for bits in (8, 16, 24):
            v, f = t.search_batch(q, psa_bits=bits, output_mode=mode)
            assert np.array_equal(f, ef), (mode, bits)
            assert np.array_equal(v, ev), (mode, bits)
    pos = np.empty(len(q), dtype=np.uint32)
    v, f = t.search_batch(q, psa_bits=16, output_mode=3, out_pos=pos)
    ko, po = oracle.psa_sort(q, 48, 64)
    assert np.array_equal(pos, po)
    assert np.array_equal(f, ef[pos]) and np.array_equal(v, ev[pos])
    # widen the key range on both sides (and touch existing keys), then search the new state
    new = np.unique(np.concatenate([g.integers(0, int(base), size=3000, dtype=np.uint64),
                                    g.integers(int(sk[-1]) + 1, 1 << 63, size=3000, dtype=np.uint64)]))
    new = np.setdiff1d(new, sk)
    upd = np.concatenate([new, sk[::97]])
    uv = W.splitmix64_mix(upd ^ np.uint64(5))
    t.update(upd, uv)
    t.publish()
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, upd, uv)
    check_range(t, cur_k)
    assert t.stats().psa_window_top == 63
    path = tmp_path / "w.hmsnap"
    t.save(path)
    t2 = BPlusTree.load(path)
    check_range(t2, cur_k)
    t2.close()
    q2 = np.concatenate([mixed_queries(cur_k, base_bits + 1), new])
    ev2, ef2 = oracle.search_sorted(cur_k, cur_v, q2)
    for mode in (1, 2):
        v, f = t.search_batch(q2, psa_bits=16, output_mode=mode)
        assert np.array_equal(f, ef2), mode
        assert np.array_equal(v, ev2), mode
    t.close()


# ---------------------------------------------------------------- PSA sort
@pytest.mark.parametrize("n", [1, 100, 4095, 4096, 4097, 300001])
@pytest.mark.parametrize("bits", [(48, 64), (0, 64), (56, 64), (40, 56), (60, 64), (3, 9)])
def test_psa_sort_bit_exact(gpu, n, bits):
    g = W.rng(n, bits[0])
    keys = g.integers(0, 1 << 63, n, dtype=np.uint64)
    if n > 10:
        keys[: n // 4] = keys[n // 2: n // 2 + n // 4]  # equal keys exercise stability
    ko, po = psa_sort(keys, *bits)
    eo, ep = oracle.psa_sort(keys, *bits)
    assert np.array_equal(po, ep) and np.array_equal(ko, eo)


@pytest.mark.parametrize("kind", ["ascending", "runs", "few_values"])
@pytest.mark.parametrize("bits", [(48, 64), (56, 64), (0, 64), (40, 56)])
def test_psa_sort_skewed_inputs(gpu, kind, bits):
    """Inputs whose waves of keys share a digit (sorted runs, a few repeated values), where the histogram
    counts a whole wave load at once, against the oracle's stable sort."""
    n = 300001
    g = W.rng(n, bits[0] + 7)
    keys = np.sort(g.integers(0, 1 << 63, n, dtype=np.uint64))
    if kind == "runs":
        keys = W.order_queries(keys, "runs4096", 5)
    elif kind == "few_values":
        keys = g.choice(np.array([3 << 50, (3 << 50) + 1, 1 << 62], dtype=np.uint64), n)
    ko, po = psa_sort(keys, *bits)
    eo, ep = oracle.psa_sort(keys, *bits)
    assert np.array_equal(po, ep) and np.array_equal(ko, eo)


@pytest.mark.parametrize("kind", ["runs", "hot"])
def test_search_psa_skewed_queries(gpu, kind):
    """PSA searches over run-sorted and hot-key query batches (the replay sort's histograms take whole
    wave loads at once) against the oracle, results in input order and in PSA order."""
    t, keys, vals, sk, sv = make(1 << 20, 64, 8, seed=61)
    if kind == "runs":
        q = W.point_queries(sk, 1 << 20, 62, order="runs4096")
    else:
        g = W.rng(63, 1)
        hot = sk[g.integers(0, len(sk), size=8)]
        q = np.concatenate([np.repeat(hot, 60000), W.point_queries(sk, 1 << 19, 64)])
        g.shuffle(q)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for mode in (1, 2):
        v, f = t.search_batch(q, psa_bits=16, output_mode=mode)
        assert np.array_equal(f, ef) and np.array_equal(v, ev), mode
    pos = np.empty(len(q), dtype=np.uint32)
    v, f = t.search_batch(q, psa_bits=16, output_mode=3, out_pos=pos)
    ko, po = oracle.psa_sort(q, 48, 64)
    assert np.array_equal(pos, po)
    assert np.array_equal(f, ef[pos]) and np.array_equal(v, ev[pos])
    t.close()


# ---------------------------------------------------------------- range
@pytest.mark.parametrize("F,fill,psa", [(64, 8, 16), (128, 8, 16), (32, 3, 16), (64, 8, 0), (128, 5, 12)])
def test_range_parity(gpu, F, fill, psa):
    """40000 ranges (with starts on keys, past the last key, 0, KEY_MAX and a zero length) against the
    oracle's leaf walk and the closed form, for several fanouts, leaf fills and PSA settings."""
    t, keys, vals, sk, sv = make(60000, F, fill, seed=41 + F, psa_bits=psa)
    starts, lens = W.range_queries(40000, 41)
    starts[:100] = sk[W.rng(1, 1).integers(0, len(sk), 100)]
    starts[100] = sk[-1] + U64(1)
    starts[101] = 0
    starts[102] = KEY_MAX
    lens[103] = 0
    L = oracle.build_bfs(sk, sv, F, fill)
    eo, ek, ev, ec = oracle.range_walk(L, starts, lens=lens)
    so, sk2, sv2, sc = oracle.range_sorted(sk, sv, starts, lens=lens)
    assert np.array_equal(ec, sc)
    off, k, v, c = t.range_batch(starts, lens)
    assert np.array_equal(off, eo) and np.array_equal(c, ec)
    for i in range(len(c)):
        a, b = int(off[i]), int(off[i]) + int(c[i])
        assert np.array_equal(k[a:b], ek[a:b]) and np.array_equal(v[a:b], ev[a:b]), i


def test_range_reference_golden(gpu):
    g = gold("range_rbpt64.npz")
    t = BPlusTree.from_keys(g["keys"], g["vals"])
    cap = int(g["harmonia_counts"].max()) + 1
    k, v, c = t.range_between(g["lo"], g["hi"], cap)
    assert np.array_equal(c, g["harmonia_counts"].astype(np.uint32))
    got = np.concatenate([k[i * cap:i * cap + c[i]] for i in range(len(c))])
    gotv = np.concatenate([v[i * cap:i * cap + c[i]] for i in range(len(c))])
    assert np.array_equal(got, g["harmonia_keys"]) and np.array_equal(gotv, g["harmonia_vals"])
    off, k, v, c = t.range_batch(g["starts"], g["lens"])
    assert np.array_equal(c, g["len_counts"])
    got = np.concatenate([k[int(off[i]):int(off[i]) + c[i]] for i in range(len(c))])
    assert np.array_equal(got, g["len_keys"])
    assert t.range_search(int(g["lo"][0]), int(g["hi"][0])) == list(
        zip(g["harmonia_keys"][: g["harmonia_counts"][0]].tolist(), g["harmonia_vals"][: g["harmonia_counts"][0]].tolist()))


# ---------------------------------------------------------------- update
def full_check(t, sk, sv, seed):
    q = np.concatenate([sk, W.gen_misses(len(sk), 5000, seed) ^ U64(1 << 40), edge(sk)])
    ev, ef = oracle.search_sorted(sk, sv, q)
    v, f = t.search_batch(q)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)
    e = t.export()
    s = e["stats"]
    lay = oracle.Layout(e["inner"], e["prefix"], e["leaves"], s.height, s.fanout,
                        np.array(s.level_start[:], dtype=np.uint32))
    assert oracle.check_layout(lay) == 0
    assert s.n_keys == len(sk)
    lk = e["leaves"][:, :8].ravel()
    assert np.array_equal(lk[lk != KEY_MAX], sk)  # the leaf chain holds exactly the key set, in order
    starts, lens = W.range_queries(3000, seed)
    off, k, vv, c = t.range_batch(starts, lens)
    eo, ek, evv, ec = oracle.range_sorted(sk, sv, starts, lens=lens)
    assert np.array_equal(c, ec)
    m = np.concatenate([np.arange(int(off[i]), int(off[i]) + int(c[i])) for i in range(len(c))]).astype(np.int64)
    assert np.array_equal(k[m], ek[m]) and np.array_equal(vv[m], evv[m])


@pytest.mark.parametrize("fill", [6, 8])
def test_update_in_place_and_relayout(gpu, fill):
    t, keys, vals, sk, sv = make(80000, 64, fill, seed=51)
    g = W.rng(51, 7)
    cur_k, cur_v = sk, sv
    rebuilt_any = False
    for step in range(4):
        new = W.gen_misses(len(keys), 400 * (step + 1), 51 + step) ^ U64(step << 50)
        new = new[~np.isin(new, cur_k)]
        old = cur_k[g.integers(0, len(cur_k), 300)]
        bk = np.concatenate([new, old, old[:50]])  # repeated keys: the later write wins
        bv = g.integers(0, 1 << 64, len(bk), dtype=np.uint64)
        g.shuffle(idx := np.arange(len(bk)))
        bk, bv = bk[idx], bv[idx]
        rebuilt_any |= t.insert_parallel(bk, bv)
        cur_k, cur_v = oracle.upsert_sorted(cur_k, cur_v, bk, bv)
        full_check(t, cur_k, cur_v, 60 + step)
    if fill == 8:
        assert rebuilt_any  # full leaves force the re-layout path


def keys_inside_leaf(cur_k, lo_i, hi_i, count, seed):
    """`count` keys strictly between cur_k[lo_i] and cur_k[hi_i] and not in cur_k."""
    g = W.rng(seed, 3)
    a, b = int(cur_k[lo_i]), int(cur_k[hi_i])
    out = np.unique(g.integers(a + 1, b, size=4 * count, dtype=np.uint64))
    out = out[~np.isin(out, cur_k)]
    assert len(out) >= count
    return out[:count]


@pytest.mark.parametrize("F", [32, 64, 128])
def test_update_node_respread(gpu, F):
    """Leaves that overflow inside a last-level node with room: that node's keys are re-spread over its
    own leaves in place (no re-layout) -- first and last leaves of nodes, two overflowing leaves in one
    node, the tree's last leaf -- with every result, the layout invariants (separator = minimum of its
    right subtree) and the leaf-window search checked; a node without room forces the re-layout."""
    fill = 6
    t, keys, vals, sk, sv = make(50000, F, fill, seed=91 + F)
    cur_k, cur_v = sk, sv
    nl = int(t.stats().n_leaves)
    g = W.rng(92, F)
    for step, leaves in enumerate(([0, F - 1, F, 3 * F + 2, 3 * F + 5], [nl - 1, 1, 2 * F + 1])):
        bk = []
        for L in leaves:
            lo = L * fill
            hi = min(lo + fill - 1, len(sk) - 1)
            if L == nl - 1:  # keys above every key land in the last leaf
                bk.append(np.arange(1, 10, dtype=np.uint64) + cur_k[-1])
            else:
                bk.append(keys_inside_leaf(sk, lo, hi, 4, 100 * step + L))
        new = np.concatenate(bk)
        old = cur_k[g.integers(0, len(cur_k), 200)]
        bkeys = np.concatenate([new, old])
        bvals = g.integers(0, 1 << 64, len(bkeys), dtype=np.uint64)
        before = t.stats()
        rebuilt = t.insert_parallel(bkeys, bvals)
        after = t.stats()
        assert not rebuilt and after.relayouts == before.relayouts
        assert after.respreads > before.respreads
        cur_k, cur_v = oracle.upsert_sorted(cur_k, cur_v, bkeys, bvals)
        full_check(t, cur_k, cur_v, 300 + step)
        q = np.concatenate([cur_k, W.gen_misses(len(keys), 20000, 5 + step)])
        ev, ef = oracle.search_sorted(cur_k, cur_v, q)
        v, f = t.search_batch(q, psa_bits=16, sorted_kernel=2)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
    # more new keys than one node's leaves can hold: the epoch re-lays every leaf out
    lo = 5 * F * fill
    crowd = keys_inside_leaf(sk, lo, lo + F * fill - 1, 3 * F, 77)
    before = t.stats()
    assert t.insert_parallel(crowd, W.splitmix64_mix(crowd))
    assert t.stats().relayouts == before.relayouts + 1
    cur_k, cur_v = oracle.upsert_sorted(cur_k, cur_v, crowd, W.splitmix64_mix(crowd))
    full_check(t, cur_k, cur_v, 400)


@pytest.mark.parametrize("size", [1, 4095, 4096, 4097, 30000, 65537, (1 << 22) + 1])
@pytest.mark.parametrize("fill", [8, 5])
def test_update_batch_sizes(gpu, size, fill):
    """Batches on all sides of the three staged-key sorts: one block (<= 4096), the merge sort (up to
    2^22) and the radix sort; each mixes new keys, existing keys and repeats (the last write wins,
    also across sort chunks), staged in two calls."""
    t, keys, vals, sk, sv = make(60000, 64, fill, seed=71 + size)
    g = W.rng(71, size)
    n_new = (size + 1) // 2
    new = W.gen_misses(len(keys), n_new, 71 + size) ^ U64(3 << 52)
    new = new[~np.isin(new, sk)]
    old = sk[g.integers(0, len(sk), size - len(new))]
    bk = np.concatenate([new, old])
    g.shuffle(bk)
    bk[len(bk) // 3:len(bk) // 3 + len(bk) // 10] = bk[:len(bk) // 10]  # repeats
    bv = g.integers(0, 1 << 64, len(bk), dtype=np.uint64)
    half = len(bk) // 2
    t.update(bk[:half], bv[:half])
    t.update(bk[half:], bv[half:])
    t.publish()
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, bk, bv)
    full_check(t, cur_k, cur_v, 90 + size)


def test_update_device_buffers(gpu):
    import torch

    t, keys, vals, sk, sv = make(30000, 64, 6, seed=81)
    new = W.gen_misses(len(keys), 3000, 81)
    nv = W.splitmix64_mix(new)
    t.update(torch.from_numpy(new.view(np.int64)).to(gpu), torch.from_numpy(nv.view(np.int64)).to(gpu))
    t.publish()
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, new, nv)
    full_check(t, cur_k, cur_v, 82)
    bad = torch.from_numpy(np.array([5, KEY_MAX], dtype=U64).view(np.int64)).to(gpu)
    with pytest.raises(HarmoniaError):
        t.update(bad, bad)
    assert t.stats().epoch == 1


def test_update_epoch_snapshot(gpu):
    t, keys, vals, sk, sv = make(20000, 64, 6, seed=61)
    bk = sk[:100]
    bv = np.full(100, 7, dtype=U64)
    t.update(bk, bv)
    v, f = t.search_batch(bk)  # staged, not yet published: old snapshot
    assert np.array_equal(v, sv[:100])
    t.publish()
    v, f = t.search_batch(bk)
    assert np.all(v == 7) and np.all(f == 1)
    with pytest.raises(HarmoniaError):
        t.update(np.array([KEY_MAX], dtype=U64), np.array([1], dtype=U64))
    assert t.stats().epoch == 1


def test_update_reference_golden(gpu):
    g = gold("update_rbpt64.npz")
    for fill in (8, 6):
        t = BPlusTree.from_keys(g["base_keys"], g["base_vals"], leaf_fill=fill)
        t.insert_parallel(g["new_keys"], g["new_vals"])
        v, f = t.search_batch(g["queries"])
        assert np.array_equal(f, g["rbpt_found"])
        assert np.array_equal(v[f == 1], g["rbpt_vals"][f == 1])


def test_update_empty_tree_and_single_inserts(gpu):
    t = BPlusTree.from_keys(np.empty(0, dtype=U64), np.empty(0, dtype=U64))
    ks = []
    for i, k in enumerate([50, 10, 30, 20, 40, 60, 70, 80, 90, 5]):
        t.insert(k, k * 3)
        ks.append(k)
        sk = np.array(sorted(ks), dtype=U64)
        full_check(t, sk, sk * U64(3), i)
    assert t.search(30) == 90 and t.search(31) is None


@pytest.mark.parametrize("F", [32, 64, 128])
def test_snapshot_roundtrip(gpu, tmp_path, F):
    """save -> load gives the identical layout and answers; a corrupted file is rejected."""
    t, keys, vals, sk, sv = make(50000, F, 6, seed=91)
    new = W.gen_misses(len(keys), 7000, 91)
    t.insert_parallel(new, W.splitmix64_mix(new))  # includes a publish
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, new, W.splitmix64_mix(new))
    path = tmp_path / "tree.hmsnap"
    t.save(path)
    t2 = BPlusTree.load(path)
    e1, e2 = t.export(), t2.export()
    for k in ("inner", "prefix", "leaves"):
        assert np.array_equal(e1[k], e2[k])
    assert t2.stats().fanout == F and t2.stats().epoch == t.stats().epoch
    full_check(t2, cur_k, cur_v, 92)
    raw = bytearray(path.read_bytes())
    for off in (200, len(raw) - 3):  # a leaf byte; a prefix byte
        bad = bytearray(raw)
        bad[off] ^= 0x5A
        p2 = tmp_path / f"bad{off}.hmsnap"
        p2.write_bytes(bytes(bad))
        with pytest.raises(HarmoniaError):
            BPlusTree.load(p2)
    (tmp_path / "short.hmsnap").write_bytes(bytes(raw[:1000]))
    with pytest.raises(HarmoniaError):
        BPlusTree.load(tmp_path / "short.hmsnap")
