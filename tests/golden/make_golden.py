"""Write golden vectors from the reference's own CPU code (oracle/_ref, compiled from the reference
sources by `make -C oracle ref`).  Run in the build container, where the reference tree exists:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs only:
  search_rbpt64.npz   regular_bpt (order 64) insert + find, and Harmonia BPlusTree insert + sort + search
  search_rbpt128.npz  regular_bpt (order 128) insert + find
  range_rbpt64.npz    regular_bpt find_range(lo, hi) and Harmonia range_search(lo, hi), closed intervals,
                      plus "first L keys >= start" ranges from range_search(start, 2^63 - 2) truncated to L.
                      regular_bpt's find_range returns nothing when `start` is above every key of the
                      leaf it descends to (H/regular_bpt/regular_bpt.cpp:325-326); Harmonia's
                      range_search walks on to the next leaf.  Both outputs are stored.
  update_rbpt64.npz   regular_bpt insert of a batch of new keys into a built tree, then find
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from harmonia_amd import workloads as W  # noqa: E402

TOP = np.uint64((1 << 63) - 2)  # largest key the reference can hold (Max_Key = 2^63 - 1 is padding)


def edge_queries(sk):
    e = [0, 1, int(sk[0]), int(sk[0]) - 1 if sk[0] > 0 else 0, int(sk[-1]), int(sk[-1]) + 1, int(TOP),
         int(sk[len(sk) // 2]) + 1, int(sk[len(sk) // 2]) - 1]
    return np.array(e, dtype=np.uint64)


def search_fixture(path, n, m, seed, order, with_harmonia):
    keys, vals = W.gen_keys(n, seed)
    assert not np.any(vals == np.uint64(0xFFFFFFFFFFFFFFFF)), "value collides with the -1 sentinel"
    sk = np.sort(keys)
    q = np.concatenate([W.point_queries(sk, m, seed), edge_queries(sk)])
    R = oracle.RefRegularTree(keys, vals, order)
    rv, rf = R.find(q)
    out = dict(keys=keys, vals=vals, queries=q, rbpt_vals=rv, rbpt_found=rf)
    if with_harmonia:
        H = oracle.RefHarmoniaTree(keys, vals)
        out["harmonia_vals"] = H.search(q)
    np.savez_compressed(path, **out)


def range_fixture(path, n, seed):
    keys, vals = W.gen_keys(n, seed)
    sk = np.sort(keys)
    g = W.rng(seed, 9)
    m = 256
    lo = np.where(g.random(m) < 0.5, sk[g.integers(0, n, m)], g.integers(0, 1 << 63, m, dtype=np.uint64))
    span = g.integers(0, 1 << 52, m, dtype=np.uint64)
    hi = np.minimum(lo + span, TOP)
    R = oracle.RefRegularTree(keys, vals, 64)
    H = oracle.RefHarmoniaTree(keys, vals)
    rk, rv, rc, hk, hv, hc = [], [], [], [], [], []
    for a, b in zip(lo, hi):
        k1, v1 = R.find_range(a, b, n)
        k2, v2 = H.range_search(a, b, n)
        rk.append(k1), rv.append(v1), rc.append(len(k1))
        hk.append(k2), hv.append(v2), hc.append(len(k2))
    # length-limited form: first L keys >= start
    starts = np.concatenate([g.integers(0, 1 << 63, m, dtype=np.uint64), sk[g.integers(0, n, m)],
                             np.array([0, sk[-1], sk[-1] + np.uint64(1)], dtype=np.uint64)])
    lens = g.integers(0, 300, len(starts)).astype(np.uint32)
    lk, lv, lc = [], [], []
    for s, L in zip(starts, lens):
        k1, v1 = H.range_search(s, TOP, n)
        lk.append(k1[:L]), lv.append(v1[:L]), lc.append(min(len(k1), int(L)))
    np.savez_compressed(path, keys=keys, vals=vals, lo=lo, hi=hi,
                        rbpt_keys=np.concatenate(rk), rbpt_vals=np.concatenate(rv), rbpt_counts=np.array(rc),
                        harmonia_keys=np.concatenate(hk), harmonia_vals=np.concatenate(hv),
                        harmonia_counts=np.array(hc),
                        starts=starts, lens=lens, len_keys=np.concatenate(lk), len_vals=np.concatenate(lv),
                        len_counts=np.array(lc, dtype=np.uint32))


def update_fixture(path, n, m, seed):
    keys, vals = W.gen_keys(n + m, seed)
    base_k, base_v = keys[:n], vals[:n]
    new_k = keys[n:]
    new_v = W.splitmix64_mix(new_k ^ np.uint64(0x5555))  # values differ from the build rule
    R = oracle.RefRegularTree(base_k, base_v, 64)
    R.insert(new_k, new_v)
    sk = np.sort(keys)
    q = np.concatenate([W.point_queries(sk, 4096, seed + 1), new_k])
    rv, rf = R.find(q)
    np.savez_compressed(path, base_keys=base_k, base_vals=base_v, new_keys=new_k, new_vals=new_v, queries=q,
                        rbpt_vals=rv, rbpt_found=rf)


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree exists")
    search_fixture(os.path.join(HERE, "search_rbpt64.npz"), 4096, 4096, 101, 64, True)
    search_fixture(os.path.join(HERE, "search_rbpt128.npz"), 10000, 4096, 202, 128, False)
    range_fixture(os.path.join(HERE, "range_rbpt64.npz"), 4096, 303)
    update_fixture(os.path.join(HERE, "update_rbpt64.npz"), 4096, 1000, 404)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
