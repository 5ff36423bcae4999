"""GPU tests at the benchmark configurations' own workloads, every result checked.

* 2^20 keys, fanout 64, 2^20 unsorted queries (90% hits): bit-exact against the closed form, the
  oracle's restatements of the reference CPU searches and, where oracle/_ref is built, the reference's
  own regular_bpt find and Harmonia BPlusTree::search.
* Uneven batches on a 2^25-key tree: 600 batches, sizes log-uniform in [32, 2^20], each uniform /
  Zipf(0.99) / runs of 4096, through the hipGraph batcher at capacity 2^24 with PSA forced on and with
  the tree's automatic choice; every batch's found flags and values are checked.
* Served mix on a 2^27-key tree: a few hundred concurrent lookup / range / insert requests through the
  serving loop; every read is checked exactly against the state of the epoch it reports, and inserted
  keys are looked up once their update has completed.
"""
import numpy as np
import pytest

import oracle
from harmonia_amd import Batcher, BPlusTree, Server
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
U64 = np.uint64


def expected(q, miss_set):
    """Found flags and values of queries whose misses are exactly the members of miss_set."""
    f = ~np.isin(q, miss_set)
    return f.astype(np.uint8), np.where(f, W.splitmix64_mix(q), U64(0))


# ------------------------------------------------------------------ 2^20 keys, 2^20 queries
def test_cpu_reference_config_on_gpu(gpu):
    import torch

    seed = W.config_seed(0)
    n = 1 << 20
    keys, vals = W.gen_keys(n, seed)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    q = W.point_queries(sk, n, seed + 7)
    print(f"2^20-key config: {len(q)} queries", flush=True)
    ev, ef = oracle.search_sorted(sk, sv, q)
    assert int(ef.sum()) == round(0.9 * n)
    tree = BPlusTree.from_keys(keys, vals, fanout=64)
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    for psa in (16, 0):
        ov = torch.full((n,), 7, dtype=torch.int64, device=gpu)
        of = torch.full((n,), 7, dtype=torch.uint8, device=gpu)
        tree.search_batch(dq, out_vals=ov, out_found=of, psa_bits=psa)
        torch.cuda.synchronize()
        assert np.array_equal(of.cpu().numpy(), ef), psa
        assert np.array_equal(ov.cpu().numpy().view(U64), ev), psa
    # the CPU restatements of the reference searches over the same layout agree with the GPU
    e = tree.export()
    s = e["stats"]
    lay = oracle.Layout(e["inner"], e["prefix"], e["leaves"], s.height, s.fanout,
                        np.array(s.level_start[:], dtype=np.uint32))
    for v, f in (oracle.search_cfb(lay, q, 8), oracle.search_harmonia(lay, q, 8)):
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
    tree.close()
    if oracle.ref_available():  # the reference's own CPU code, compiled from its sources
        part = slice(0, 1 << 18)
        rv, rf = oracle.RefRegularTree(keys, vals, 64).find(q[part], nthreads=8)
        assert np.array_equal(rf, ef[part]) and np.array_equal(rv, ev[part])
        hv = oracle.RefHarmoniaTree(keys, vals).search(q[part], nthreads=8)  # -1 = not found
        hit = ef[part].astype(bool)
        assert (hv[~hit] == -1).all() and np.array_equal(hv[hit].view(U64), ev[part][hit])


# ------------------------------------------------------------------ uneven batches, 2^25 keys
@pytest.fixture(scope="module")
def uneven(gpu):
    seed = W.config_seed(3)
    keys, vals = W.gen_keys(1 << 25, seed)
    sk = np.sort(keys)
    pool_n = 1 << 23
    misses = W.gen_misses(len(sk), pool_n, seed + 1)
    uni = W.point_queries(sk, pool_n, seed + 1)
    pools = [uni, W.zipf_queries(sk, pool_n, seed + 2), W.order_queries(uni, "runs4096", seed + 3)]
    exp = [expected(p, misses) for p in pools]
    tree = BPlusTree.from_keys(keys, vals, fanout=64)
    yield tree, sk, W.splitmix64_mix(sk), pools, exp
    tree.close()


@pytest.mark.parametrize("psa_bits", [16, -1])  # PSA forced on every flush / the tree's threshold
def test_uneven_batches_through_batcher(uneven, psa_bits):
    import torch

    tree, sk, sv, pools, exp = uneven
    n_batches = 600
    g = W.rng(W.config_seed(3), 11 + psa_bits)
    sizes = W.log_uniform_sizes(n_batches, 32, 1 << 20, W.config_seed(3) + psa_bits)
    dist = g.integers(0, 3, n_batches)
    starts = [int(g.integers(0, len(pools[0]) - int(m) + 1)) for m in sizes]
    dpools = [torch.from_numpy(p.view(np.int64)).to("cuda:0") for p in pools]
    cap = 1 << 24
    b = Batcher(tree, cap, psa_bits=psa_bits)
    print(f"uneven: {n_batches} batches, {int(sizes.sum())} queries, psa_bits {psa_bits}", flush=True)
    pending, fill, checked = [], 0, 0
    sample_q, sample_v, sample_f = [], [], []
    for i in range(n_batches + 1):
        if i == n_batches or fill + int(sizes[i]) > cap:
            b.flush()
            b.sync()
            v_all, f_all = b.copy_out(0, fill)
            for j, off in pending:
                m, d, s0 = int(sizes[j]), int(dist[j]), starts[j]
                ef, ev = exp[d][0][s0:s0 + m], exp[d][1][s0:s0 + m]
                v, f = v_all[off:off + m], f_all[off:off + m]
                assert np.array_equal(f, ef), (j, m, d)
                assert np.array_equal(v, ev), (j, m, d)
                checked += 1
                if j % 7 == 0:
                    k = min(m, 256)
                    sample_q.append(pools[d][s0:s0 + k])
                    sample_v.append(v[:k])
                    sample_f.append(f[:k])
            pending, fill = [], 0
            if i == n_batches:
                break
        m, d, s0 = int(sizes[i]), int(dist[i]), starts[i]
        pending.append((i, b.add(dpools[d][s0:s0 + m])))
        fill += m
    b.close()
    assert checked == n_batches
    q = np.concatenate(sample_q)
    ev, ef = oracle.search_sorted(sk, sv, q)
    assert np.array_equal(np.concatenate(sample_f), ef) and np.array_equal(np.concatenate(sample_v), ev)


# ------------------------------------------------------------------ served mix, 2^27 keys
def test_served_mix_2p27(gpu):
    seed = W.config_seed(4)
    n = 1 << 27
    keys, vals = W.gen_keys(n, seed)
    tree = BPlusTree.from_keys(keys, vals, fanout=64, leaf_fill=5)
    del vals
    sk = np.sort(keys)
    del keys
    g = W.rng(seed, 1)
    n_req = 320
    kinds = g.choice(3, n_req, p=[0.85, 0.10, 0.05])
    upd_sizes = np.round(np.exp(g.uniform(0, np.log(4096), n_req))).astype(np.int64)
    fresh = W.gen_misses(n, int(upd_sizes[kinds == 2].sum()) + 200000, seed)  # never in the initial set
    ins_pool, miss_pool = fresh[:-200000], fresh[-200000:]
    ins_vals = W.splitmix64_mix(ins_pool ^ U64(0x55))
    srv = Server(tree, max_inflight=4)
    reqs, u0 = [], 0
    print(f"served mix: {n_req} requests on 2^27 keys", flush=True)
    for i in range(n_req):
        if kinds[i] == 0:
            m = int(np.round(np.exp(g.uniform(0, np.log(65536)))))
            q = np.concatenate([sk[g.integers(0, n, m - m // 10)], miss_pool[g.integers(0, len(miss_pool), m // 10)],
                                ins_pool[g.integers(0, max(u0, 1), min(u0, 64))]])
            reqs.append(("lookup", q, srv.lookup(q)))
        elif kinds[i] == 1:
            st, ln = W.range_queries(1, seed + i)
            reqs.append(("range", (st, ln), srv.range(st, ln)))
        else:
            m = int(upd_sizes[i])
            part = slice(u0, u0 + m)
            u0 += m
            reqs.append(("update", part, srv.update(ins_pool[part], ins_vals[part])))
    done = []
    for kind, payload, tk in reqs:
        st, ep, out = srv.wait(tk)
        assert st == 0, kind
        done.append((kind, payload, ep, out))
    # epoch at which each inserted key became visible
    vis = np.full(u0, np.iinfo(np.int64).max, dtype=np.int64)
    for kind, payload, ep, out in done:
        if kind == "update":
            vis[payload] = ep
    order = np.argsort(ins_pool[:u0])
    ins_sorted, ins_vis, ins_v = ins_pool[:u0][order], vis[order], ins_vals[:u0][order]
    for kind, payload, ep, out in done:
        if kind == "lookup":
            q = payload
            v, f = out
            if u0:
                i = np.minimum(np.searchsorted(ins_sorted, q), u0 - 1)
                is_ins = ins_sorted[i] == q
                ins_ok = is_ins & (ins_vis[i] <= ep)
                ins_val = ins_v[i]
            else:
                is_ins = ins_ok = np.zeros(len(q), bool)
                ins_val = np.zeros(len(q), U64)
            in_init = ~is_ins & ~np.isin(q, miss_pool)
            ef = in_init | ins_ok
            evv = np.where(in_init, W.splitmix64_mix(q), np.where(ins_ok, ins_val, U64(0)))
            assert np.array_equal(f.astype(bool), ef), ep
            assert np.array_equal(v, evv), ep
        elif kind == "range":
            (st, ln), (off, k, v, c) = payload, out
            lb = int(np.searchsorted(sk, st[0]))
            L = int(ln[0])
            live = ins_sorted[ins_vis <= ep]
            lvv = ins_v[ins_vis <= ep]
            j = int(np.searchsorted(live, st[0]))
            mk = np.concatenate([sk[lb:lb + L], live[j:j + L]])
            mv = np.concatenate([W.splitmix64_mix(sk[lb:lb + L]), lvv[j:j + L]])
            o = np.argsort(mk, kind="stable")[:L]
            assert int(c[0]) == len(o)
            assert np.array_equal(k[:len(o)], mk[o]) and np.array_equal(v[:len(o)], mv[o])
    # every inserted key is found with its value once its update has completed
    for a in range(0, u0, 65536):
        qs = ins_pool[a:a + 65536]
        st, ep, (v, f) = srv.wait(srv.lookup(qs))
        assert st == 0 and f.all() and np.array_equal(v, ins_vals[a:a + 65536])
    srv.drain()
    s = srv.stats()
    assert s["updates"] == int((kinds == 2).sum()) and s["epochs"] >= 1
    srv.close()
    tree.close()
