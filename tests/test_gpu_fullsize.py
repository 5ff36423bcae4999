"""Full-size checks on the 2^26-key, fanout-64 tree with 2^24 queries: size-independent properties
over every result (found <=> the query is in the key set; value = splitmix64_mix(query)), plus a
bit-exact sample against the oracle, for both query orders and PSA on and off."""
import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu

N_KEYS = 1 << 26
N_Q = 1 << 24


@pytest.fixture(scope="module")
def big(gpu):
    import torch

    keys, vals = W.gen_keys(N_KEYS, W.config_seed(1))
    dk = torch.from_numpy(keys.view(np.int64)).to(gpu)
    dv = torch.from_numpy(vals.view(np.int64)).to(gpu)
    t = BPlusTree.from_keys(dk, dv, fanout=64, leaf_fill=8)
    del dk, dv
    sk = np.sort(keys)
    return t, sk


@pytest.mark.parametrize("order", ["unsorted", "runs4096"])
def test_fullsize_search_properties(big, gpu, order):
    import torch

    t, sk = big
    q = W.point_queries(sk, N_Q, W.config_seed(1) + 7, order=order)
    hit = np.isin(q, sk, assume_unique=False)
    assert hit.sum() == round(0.9 * N_Q)
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    dv = torch.empty(N_Q, dtype=torch.int64, device=gpu)
    df = torch.empty(N_Q, dtype=torch.uint8, device=gpu)
    sample = W.rng(3, 3).integers(0, N_Q, 1 << 16)
    ev, ef = oracle.search_sorted(sk, W.splitmix64_mix(sk), q[sample])
    for psa in (0, 16):
        print(f"fullsize order={order} psa={psa}", flush=True)
        t.search_batch(dq, out_vals=dv, out_found=df, psa_bits=psa)
        torch.cuda.synchronize()
        f = df.cpu().numpy()
        v = dv.cpu().numpy().view(np.uint64)
        assert np.array_equal(f.astype(bool), hit)
        assert np.array_equal(v[hit], W.splitmix64_mix(q[hit]))
        assert np.all(v[~hit] == 0)
        assert np.array_equal(f[sample], ef) and np.array_equal(v[sample], ev)


def test_fullsize_layout_shape(big):
    t, sk = big
    s = t.stats()
    h, ni, nl, ls = oracle.plan(N_KEYS, 64, 8)
    assert (s.height, s.n_inner, s.n_leaves) == (h, ni, nl)
    assert s.n_leaves == N_KEYS // 8
    e = t.export()
    lk = e["leaves"][:, :8].ravel()
    assert np.array_equal(lk, sk)  # full leaves: the leaf chain is exactly the sorted key set
