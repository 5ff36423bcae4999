"""Seeded synthetic workloads for the tree (keys, values, query mixes, ranges, update batches).

Keys are unique by construction: key_i = perm63(base + i), where perm63 is a bijection of
[0, 2^63) (multiply by an odd constant and xor-shift, each invertible modulo 2^63), so no
duplicate check is needed and keys look uniform over [0, 2^63).  2^63 - 1 (the reference's
Max_Key = LLONG_MAX, H/src/bpt.h:25) is never produced.  Misses are perm63 of inputs outside
[base, base + n), hence absent from the key set.  The value of a key is splitmix64_mix(key) =
the first output of Vigna's splitmix64 stream whose state is the key.
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
MASK63 = np.uint64((1 << 63) - 1)
RESERVED63 = (1 << 63) - 1
SEED_BASE = 0x4841524D4F4E4941  # "HARMONIA"


def _mix(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def splitmix64_mix(keys):
    """value(key): first splitmix64 output for state = key (vectorised, uint64)."""
    keys = np.asarray(keys, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(keys + GAMMA)


def splitmix64_stream(seed, n):
    """n consecutive outputs of splitmix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        st = np.uint64(seed) + GAMMA * np.arange(1, n + 1, dtype=np.uint64)
        return _mix(st)


def perm63(x, salt):
    """A seeded bijection of [0, 2^63)."""
    x = np.asarray(x, dtype=np.uint64) & MASK63
    salt = np.uint64(salt) & MASK63
    with np.errstate(over="ignore"):
        x = (x ^ salt) & MASK63
        x = (x * np.uint64(0x9E3779B97F4A7C15 | 1)) & MASK63
        x ^= x >> np.uint64(31)
        x = (x * M1) & MASK63
        x ^= x >> np.uint64(29)
        x = (x * M2) & MASK63
        x ^= x >> np.uint64(32)
    return x


def config_seed(config_index):
    return SEED_BASE ^ int(config_index)


def _salts(seed):
    s = splitmix64_stream(seed, 3)
    return int(s[0] & MASK63), int(s[1] >> np.uint64(2)), int(s[2])


def gen_keys(n, seed):
    """n unique uint64 keys in [0, 2^63 - 1), pseudo-random order; values = splitmix64_mix."""
    salt, base, _ = _salts(seed)
    idx = np.arange(n + 1, dtype=np.uint64) + np.uint64(base)
    k = perm63(idx, salt)
    bad = np.nonzero(k[:n] == np.uint64(RESERVED63))[0]
    keys = k[:n].copy()
    if bad.size:  # at most one input maps to 2^63-1; use the spare input base + n
        keys[bad] = k[n]
    return keys, splitmix64_mix(keys)


def gen_misses(n_keys, m, seed):
    """m keys guaranteed absent from gen_keys(n_keys, seed)."""
    salt, base, _ = _salts(seed)
    idx = np.arange(m + 2, dtype=np.uint64) + np.uint64(base) + np.uint64(n_keys + 1)
    k = perm63(idx, salt)
    k = k[k != np.uint64(RESERVED63)]
    return k[:m].copy()


def rng(seed, stream):
    return np.random.Generator(np.random.PCG64([int(seed) & ((1 << 64) - 1), int(stream)]))


def point_queries(keys, m, seed, hit_frac=0.9, order="unsorted"):
    """m point queries: round(hit_frac*m) uniform hits over `keys`, the rest misses.

    order: "unsorted" (shuffled), "sorted", or "runs4096" (sorted, cut into runs of 4096
    keys, runs shuffled)."""
    n = len(keys)
    g = rng(seed, 1)
    n_hit = int(round(hit_frac * m)) if n > 0 else 0
    hits = keys[g.integers(0, n, size=n_hit)] if n_hit else np.empty(0, dtype=np.uint64)
    misses = gen_misses(n, m - n_hit, seed)
    q = np.concatenate([hits, misses]).astype(np.uint64)
    g.shuffle(q)
    return order_queries(q, order, seed)


def order_queries(q, order, seed, run=4096):
    if order == "unsorted":
        return q
    s = np.sort(q)
    if order == "sorted":
        return s
    if order == "runs4096":
        nr = (len(s) + run - 1) // run
        perm = rng(seed, 2).permutation(nr)
        return np.concatenate([s[i * run:(i + 1) * run] for i in perm]) if nr else s
    raise ValueError(order)


def zipf_ranks(n, m, s, seed):
    """m ranks in [1, n] drawn from Zipf(s) by rejection-inversion (Hoermann & Derflinger 1996)."""
    g = rng(seed, 3)
    one_minus = 1.0 - s

    def h(x):
        return np.exp(-s * np.log(x))

    def helper1(x):
        return np.where(np.abs(x) > 1e-8, np.log1p(x) / np.where(x == 0, 1, x), 1 - x / 2 + x * x / 3)

    def helper2(x):
        return np.where(np.abs(x) > 1e-8, np.expm1(x) / np.where(x == 0, 1, x), 1 + x / 2 + x * x / 6)

    def H(x):
        lx = np.log(x)
        return helper2(one_minus * lx) * lx

    def Hinv(x):
        t = x * one_minus
        t = np.maximum(t, -1 + 1e-16)
        return np.exp(helper1(t) * x)

    hx1 = H(np.float64(1.5)) - 1.0
    hn = H(np.float64(n + 0.5))
    sc = 2.0 - Hinv(H(np.float64(2.5)) - h(np.float64(2.0)))
    out = np.empty(m, dtype=np.int64)
    todo = np.arange(m)
    while todo.size:
        u = hn + g.random(todo.size) * (hx1 - hn)
        x = Hinv(u)
        k = np.clip(np.floor(x + 0.5), 1, n)
        ok = (k - x <= sc) | (u >= H(k + 0.5) - h(k))
        out[todo[ok]] = k[ok].astype(np.int64)
        todo = todo[~ok]
    return out


def zipf_queries(sorted_keys, m, seed, s=0.99):
    """m queries: key of Zipf(s) rank over the sorted key order (rank 1 = smallest key)."""
    r = zipf_ranks(len(sorted_keys), m, s, seed)
    return sorted_keys[r - 1]


def range_queries(m, seed, lo=16, hi=256):
    """m ranges: start uniform in [0, 2^63), length uniform in [lo, hi]."""
    g = rng(seed, 4)
    starts = g.integers(0, 1 << 63, size=m, dtype=np.uint64)
    lens = g.integers(lo, hi + 1, size=m).astype(np.uint32)
    return starts, lens


def log_uniform_sizes(count, lo, hi, seed, stream=5):
    g = rng(seed, stream)
    return np.round(np.exp(g.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)
