"""Seeded workload generators: uniqueness, domain, determinism and the query-mix shapes."""
import numpy as np

from harmonia_amd import workloads as W

U64 = np.uint64


def test_keys_unique_in_domain_and_deterministic():
    k, v = W.gen_keys(200000, 5)
    assert len(np.unique(k)) == len(k)
    assert np.all(k < U64((1 << 63) - 1))
    k2, v2 = W.gen_keys(200000, 5)
    assert np.array_equal(k, k2) and np.array_equal(v, v2)
    assert np.array_equal(v, W.splitmix64_mix(k))
    # roughly uniform over [0, 2^63): each eighth holds about 1/8 of the keys
    h = np.bincount((k >> U64(60)).astype(np.int64), minlength=8)
    assert np.all(np.abs(h - len(k) / 8) < 0.02 * len(k))


def test_perm63_is_a_bijection_on_a_block():
    x = np.arange(1 << 16, dtype=U64) + U64(12345)
    y = W.perm63(x, 999)
    assert len(np.unique(y)) == len(y) and np.all(y < U64(1 << 63))


def test_misses_absent():
    k, _ = W.gen_keys(100000, 8)
    m = W.gen_misses(100000, 50000, 8)
    assert len(m) == 50000 and not np.any(np.isin(m, k))
    assert np.all(m != U64((1 << 63) - 1))


def test_point_query_mix():
    k, _ = W.gen_keys(100000, 9)
    sk = np.sort(k)
    q = W.point_queries(sk, 100000, 9)
    hit = np.isin(q, sk)
    assert hit.sum() == 90000
    r = W.point_queries(sk, 100000, 9, order="runs4096")
    assert np.array_equal(np.sort(r), np.sort(q))
    # sorted runs of 4096 in shuffled order: at most (#runs - 1) descents, and some exist
    descents = int(np.sum(r[1:] < r[:-1]))
    assert 0 < descents <= (len(r) + 4095) // 4096 - 1


def test_zipf_rank_frequencies():
    n, m, s = 1 << 20, 400000, 0.99
    r = W.zipf_ranks(n, m, s, 3)
    assert r.min() >= 1 and r.max() <= n
    hn = np.sum(1.0 / np.arange(1, n + 1) ** s)
    for rank in (1, 2, 10):
        expect = m / rank ** s / hn
        got = np.sum(r == rank)
        assert abs(got - expect) < 5 * np.sqrt(expect) + 0.01 * expect


def test_ranges():
    s, L = W.range_queries(10000, 4)
    assert L.min() >= 16 and L.max() <= 256 and s.dtype == np.uint64
