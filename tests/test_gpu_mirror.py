"""The C++ mirror of the reference's BPlusTree interface (include/harmonia.hpp) on the GPU: a driver
built against it (harmonia_amd/hm_mirror_example) inserts, builds, searches, range-searches and
upserts; everything it returns is compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
U64 = np.uint64
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "harmonia_amd", "hm_mirror_example")


def stream63(seed, n):
    return W.splitmix64_stream(seed, n) >> U64(1)


def test_cpp_mirror_on_gpu(gpu, tmp_path):
    assert os.path.exists(EXE), "make -C harmonia_amd builds hm_mirror_example"
    n, seed = 100000, 5
    out = tmp_path / "res.bin"
    print(f"mirror driver: {n} keys, seed {seed}", flush=True)
    subprocess.run([EXE, str(n), str(seed), str(out)], check=True, timeout=120)
    r = np.fromfile(out, dtype=U64)
    keys = stream63(seed, n)
    assert len(np.unique(keys)) == n
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    q = np.concatenate([keys, stream63(seed + 1, n)])
    pos = 0

    def take(m):
        nonlocal pos
        a = r[pos:pos + m]
        pos += m
        return a

    ev, ef = oracle.search_sorted(sk, sv, q)
    got = take(2 * len(q)).reshape(-1, 2)
    assert np.array_equal(got[:, 0], ef) and np.array_equal(got[:, 1], ev)
    got = take(2 * 1000).reshape(-1, 2)
    assert np.array_equal(got[:, 0], ef[:1000]) and np.array_equal(got[:, 1], ev[:1000])
    for i in range(200):
        lo, hi = q[i], q[i] + U64(1 << 52)
        a, b = np.searchsorted(sk, lo, "left"), np.searchsorted(sk, hi, "right")
        c = int(take(1)[0])
        assert c == b - a, i
        kv = take(2 * c).reshape(-1, 2)
        assert np.array_equal(kv[:, 0], sk[a:b]) and np.array_equal(kv[:, 1], sv[a:b])
    nk = stream63(seed + 2, n // 4)
    k2, v2 = oracle.upsert_sorted(sk, sv, nk, nk ^ U64(0x5BD1E995))
    q2 = np.concatenate([q, nk])
    ev, ef = oracle.search_sorted(k2, v2, q2)
    got = take(2 * len(q2)).reshape(-1, 2)
    assert np.array_equal(got[:, 0], ef) and np.array_equal(got[:, 1], ev)
    assert pos == len(r)
