"""The hipGraph-captured batched pipeline: uneven batches packed into one replay, every replay checked
against the oracle with eager work and synchronisation between replays, and re-capture after the
tree's layout changes."""
import numpy as np
import pytest

import oracle
from harmonia_amd import Batcher, BPlusTree, HarmoniaError
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
U64 = np.uint64


def make(n, seed, **kw):
    keys, vals = W.gen_keys(n, seed)
    t = BPlusTree.from_keys(keys, vals, **kw)
    sk = np.sort(keys)
    return t, sk, W.splitmix64_mix(sk)


def uneven_batches(sk, seed, count, lo=1, hi=5000):
    sizes = W.log_uniform_sizes(count, lo, hi, seed)
    return [W.point_queries(sk, int(m), seed + i) for i, m in enumerate(sizes)]


@pytest.mark.parametrize("psa_bits,mode", [(-1, 0), (0, 0), (16, 1), (16, 2), (24, 1)])
def test_batcher_replays_match_oracle(gpu, psa_bits, mode):
    t, sk, sv = make(200000, 7, sort_threshold=4096)
    b = Batcher(t, 1 << 16, psa_bits=psa_bits, output_mode=mode)
    for replay in range(4):
        batches = uneven_batches(sk, 100 * replay + 3, 30)
        offs = []
        total = 0
        for q in batches:
            if total + len(q) > b.capacity:
                break
            offs.append((b.add(q), q))
            total += len(q)
        b.flush()
        # eager work and a synchronisation between replays
        v_eager, f_eager = t.search_batch(np.concatenate([q for _, q in offs]))
        b.sync()
        got_v, got_f = b.copy_out(0, total)
        assert np.array_equal(got_v, v_eager) and np.array_equal(got_f, f_eager), replay
        for off, q in offs:
            ev, ef = oracle.search_sorted(sk, sv, q)
            v, f = b.copy_out(off, len(q))
            assert np.array_equal(f, ef) and np.array_equal(v, ev), (replay, off)
    b.close()


def test_batcher_full_and_overflow(gpu):
    t, sk, sv = make(50000, 8)
    b = Batcher(t, 1000)
    q = W.point_queries(sk, 1000, 1)
    assert b.add(q) == 0
    with pytest.raises(HarmoniaError):
        b.add(q[:1])
    b.flush()
    v, f = b.copy_out(0, 1000)
    ev, ef = oracle.search_sorted(sk, sv, q)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)
    b.flush()  # empty flush is a no-op
    b.close()


def test_batcher_recaptures_after_relayout(gpu):
    t, sk, sv = make(30000, 9, leaf_fill=8, sort_threshold=1)
    b = Batcher(t, 1 << 14)
    q = W.point_queries(sk, 5000, 2)
    b.add(q)
    b.flush()
    b.sync()
    new = W.gen_misses(30000, 3000, 9)
    nv = W.splitmix64_mix(new ^ U64(99))
    assert t.insert_parallel(new, nv)  # full leaves: re-layout, new device arrays
    k2, v2 = oracle.upsert_sorted(sk, sv, new, nv)
    q2 = np.concatenate([q, new[:2000]])
    b.add(q2)
    b.flush()
    v, f = b.copy_out(0, len(q2))
    ev, ef = oracle.search_sorted(k2, v2, q2)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)
    b.close()


def test_autotune_keeps_results_exact(gpu):
    keys, vals = W.gen_keys(300000, 12)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    sample = W.point_queries(sk, 1 << 15, 13)
    t, r = BPlusTree.from_keys_tuned(keys, vals, sample, quick=True)
    assert r["fanout"] in (32, 64, 128) and r["trials"] > 50 and r["best_ms"] > 0
    assert t.stats().fanout == r["fanout"]
    for order in ("unsorted", "runs4096"):
        q = W.point_queries(sk, 200000, 14, order=order)
        ev, ef = oracle.search_sorted(sk, sv, q)
        v, f = t.search_batch(q)
        assert np.array_equal(f, ef) and np.array_equal(v, ev)
    r2 = t.autotune(q[:4096], quick=True)
    assert r2["fanout"] == r["fanout"]
    v, f = t.search_batch(q)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)


def test_batcher_device_and_host_batches(gpu):
    """Device batches (gathered by the flush, pieces of 2^16), host batches (staged), mixed in one
    flush, several flushes back to back before one sync."""
    import torch

    t, sk, sv = make(300000, 12)
    b = Batcher(t, 1 << 19)
    rounds = []
    for r in range(3):
        sizes = [70000, 5, 131072 + 17, 3000, 1, 40000]
        offs = []
        for i, m in enumerate(sizes):
            q = W.point_queries(sk, m, 1000 * r + i)
            src = torch.from_numpy(q.view(np.int64)).to(gpu) if i % 2 == 0 else q
            offs.append((b.add(src), q))
        b.flush()
        out = [(off, q, torch.empty(len(q), dtype=torch.int64, device=gpu),
                torch.empty(len(q), dtype=torch.uint8, device=gpu)) for off, q in offs]
        rounds.append(out)
        # device-to-device copies in stream order, before the next round's flush reuses the buffers
        for off, q, v, f in out:
            b.copy_out(off, len(q), v, f)
    b.sync()
    for out in rounds:
        for off, q, v, f in out:
            ev, ef = oracle.search_sorted(sk, sv, q)
            assert np.array_equal(f.cpu().numpy(), ef) and np.array_equal(v.cpu().numpy().view(np.uint64), ev), off
    b.close()
