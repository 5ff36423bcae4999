"""GPU parity of the comparator baselines (pointer-child trees, hm_baseline_search): same results as
the oracle for every fanout, leaf fill, group size and node-search kind, also after update epochs
that re-lay out the tree (the child array follows the layout)."""
import numpy as np
import pytest

import oracle
from harmonia_amd import HarmoniaError
from harmonia_amd import workloads as W

from test_gpu_parity import U64, make, queries

pytestmark = pytest.mark.gpu


def run_baseline(gpu, t, q, kind, group):
    import torch

    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    dv = torch.full((len(q),), -1, dtype=torch.int64, device=gpu)
    df = torch.full((len(q),), 7, dtype=torch.uint8, device=gpu)
    s = torch.cuda.Stream()
    t.search_baseline(dq, dv, df, kind=kind, group_size=group, stream=s)
    s.synchronize()
    return dv.cpu().numpy().view(U64), df.cpu().numpy()


@pytest.mark.parametrize("n,F,fill", [(0, 64, 8), (1, 32, 8), (513, 64, 8), (100003, 32, 5), (100003, 64, 8),
                                      (250000, 128, 6)])
def test_baseline_matches_oracle(gpu, n, F, fill):
    print(f"baseline n={n} F={F} fill={fill}", flush=True)
    t, keys, vals, sk, sv = make(n, F, fill, seed=n + 3 * F)
    q = queries(sk, 20000, n + 5)
    ev, ef = oracle.search_sorted(sk, sv, q)
    for kind in ("pointer", "linear", "prefix"):
        for g in (1, 2, 4, 8):
            v, f = run_baseline(gpu, t, q, kind, g)
            assert np.array_equal(f, ef), (kind, g)
            assert np.array_equal(v, ev), (kind, g)
    t.close()


@pytest.mark.parametrize("m", [1, 31, 32, 33, 255, 257])
def test_baseline_ragged_batches(gpu, m):
    """Batch sizes that leave partial groups and partial blocks."""
    t, keys, vals, sk, sv = make(40000, 64, 8, seed=77)
    q = queries(sk, m, 78)[:m]
    ev, ef = oracle.search_sorted(sk, sv, q)
    for g in (1, 8):
        v, f = run_baseline(gpu, t, q, "linear", g)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)


def test_baseline_follows_update_epochs(gpu):
    t, keys, vals, sk, sv = make(60000, 64, 8, seed=91)  # full leaves: the first insert re-lays out
    cur_k, cur_v = sk, sv
    for step in range(3):
        new = W.gen_misses(len(keys), 3000, 92 + step) ^ U64(step << 52)
        new = new[~np.isin(new, cur_k)]
        old = cur_k[:: max(1, len(cur_k) // 500)]
        bk = np.concatenate([new, old])
        bv = W.rng(93, step).integers(0, 1 << 64, len(bk), dtype=np.uint64)
        t.insert_parallel(bk, bv)
        cur_k, cur_v = oracle.upsert_sorted(cur_k, cur_v, bk, bv)
        q = queries(cur_k, 30000, 94 + step)
        ev, ef = oracle.search_sorted(cur_k, cur_v, q)
        for kind in ("pointer", "linear", "prefix"):
            v, f = run_baseline(gpu, t, q, kind, 4)
            assert np.array_equal(f, ef) and np.array_equal(v, ev), (step, kind)


def test_baseline_rejects_bad_arguments(gpu):
    import torch

    t, keys, vals, sk, sv = make(1000, 64, 8, seed=5)
    q = queries(sk, 100, 6)
    hv, hf = np.empty(len(q), np.uint64), np.empty(len(q), np.uint8)
    with pytest.raises(HarmoniaError):
        t.search_baseline(q, hv, hf)  # host buffers
    dq = torch.from_numpy(q.view(np.int64)).to(gpu)
    dv = torch.empty(len(q), dtype=torch.int64, device=gpu)
    df = torch.empty(len(q), dtype=torch.uint8, device=gpu)
    with pytest.raises(HarmoniaError):
        t.search_baseline(dq, dv, df, group_size=3)
    assert t._lib.hm_baseline_search(t._handle(), 9, 1, dq.data_ptr(), len(q), dv.data_ptr(), df.data_ptr(),
                                     None) != 0  # unknown kind
