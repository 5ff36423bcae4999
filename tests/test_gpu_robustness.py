"""Failure paths that must leave the tree consistent:

* an epoch whose re-layout exceeds the tree's device-memory budget fails with HM_EBUDGET and drops its
  staged batch, so the next epoch does not apply it (directly through hm_epoch_publish, and through the
  serving loop, whose clients are told the update failed);
* a snapshot whose header counts disagree with the file is rejected without sizing anything from them;
* the autotuner's sort threshold is at most the sample size whenever PSA won at the sample size, also for
  a sample that is not a power of two."""
import numpy as np
import pytest

import oracle
from harmonia_amd import BPlusTree, HarmoniaError, Server
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
HM_EBUDGET = -3


def tight_tree(n, seed):
    """A fanout-64 tree at fill 8 (every leaf and last-level node full) with a budget that admits update
    scratch but not a second copy of the leaves (a re-layout's standby buffers)."""
    keys, vals = W.gen_keys(n, seed)
    probe = BPlusTree.from_keys(keys, vals, fanout=64, leaf_fill=8)
    base = probe.stats().bytes_device
    probe.close()
    t = BPlusTree.from_keys(keys, vals, fanout=64, leaf_fill=8, mem_budget=base + (4 << 20))
    return t, keys, vals


def check_state(t, sk, sv, q):
    ev, ef = oracle.search_sorted(sk, sv, q)
    v, f = t.search_batch(q)
    assert np.array_equal(f, ef) and np.array_equal(v, ev)


def test_failed_publish_drops_its_batch(gpu):
    t, keys, vals = tight_tree(1 << 20, 301)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    ep0 = t.stats().epoch
    new = W.gen_misses(len(keys), 4096, 302)  # new keys: full leaves in full nodes overflow -> re-layout
    t.update(new, W.splitmix64_mix(new))
    with pytest.raises(HarmoniaError) as e:
        t.publish()
    assert e.value.code == HM_EBUDGET
    st = t.stats()
    assert st.epoch == ep0 and st.n_keys == len(keys) and st.relayouts == 0
    # the next epoch (values of existing keys only: in place) applies its own batch and nothing else
    old = sk[::257]
    nv = old ^ np.uint64(0x5BD1E995)
    t.update(old, nv)
    assert t.publish() is False
    assert t.stats().epoch == ep0 + 1 and t.stats().n_keys == len(keys)
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, old, nv)
    check_state(t, cur_k, cur_v, np.concatenate([new, old, W.point_queries(sk, 50000, 303)]))
    t.close()


def test_failed_epoch_in_server_is_not_applied_later(gpu):
    t, keys, vals = tight_tree(1 << 20, 311)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    sv_ = Server(t, max_inflight=2)
    try:
        new = W.gen_misses(len(keys), 4096, 312)
        st, _, _ = sv_.wait(sv_.update(new, W.splitmix64_mix(new)))
        assert st == HM_EBUDGET  # the client is told the update failed
        old = sk[::331]
        nv = old ^ np.uint64(0x1234567)
        st, ep, _ = sv_.wait(sv_.update(old, nv))
        assert st == 0
        st, ep2, (v, f) = sv_.wait(sv_.lookup(np.concatenate([new, old])))
        assert st == 0 and ep2 >= ep
        assert not f[:len(new)].any() and not v[:len(new)].any()
        assert f[len(new):].all() and np.array_equal(v[len(new):], nv)
    finally:
        sv_.close()
    cur_k, cur_v = oracle.upsert_sorted(sk, sv, old, nv)
    check_state(t, cur_k, cur_v, np.concatenate([new, W.point_queries(sk, 50000, 313)]))
    t.close()


def test_snapshot_header_counts_checked(gpu, tmp_path):
    keys, vals = W.gen_keys(30000, 321)
    t = BPlusTree.from_keys(keys, vals, fanout=64, leaf_fill=6)
    path = tmp_path / "t.hmsnap"
    t.save(path)
    raw = bytearray(path.read_bytes())
    # header: magic[8], version, fanout, leaf_fill, node_slots, height, n_inner (u32), n_leaves, n_keys (u64)
    for off, val in ((32, 1 << 40), (32, 0xFFFFFFFF), (28, 0x7FFFFFFF), (20, 64)):
        bad = bytearray(raw)
        width = 8 if off == 32 else 4
        bad[off:off + width] = int(val).to_bytes(width, "little")
        p = tmp_path / f"bad_{off}_{val}.hmsnap"
        p.write_bytes(bytes(bad))
        with pytest.raises(HarmoniaError):
            BPlusTree.load(p)
    t2 = BPlusTree.load(path)  # the good file still loads
    assert t2.stats().n_keys == len(keys)
    t.close()
    t2.close()


def test_tuner_threshold_with_non_power_of_two_sample(gpu):
    keys, vals = W.gen_keys(1 << 20, 331)
    sk = np.sort(keys)
    sample = W.point_queries(sk, (1 << 18) - 1, 332)
    t = BPlusTree.from_keys(keys, vals, fanout=64)
    r = t.autotune(sample, quick=True)
    if r["psa_ms"] < r["direct_ms"]:
        assert r["sort_threshold"] <= len(sample)
    else:
        assert r["sort_threshold"] >= 1024
    check_state(t, sk, W.splitmix64_mix(sk), W.point_queries(sk, 100000, 333))
    t.close()
