"""bench.py's command-line contract and its byte models, on the CPU (no GPU is touched)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def bench():
    """bench.py as a module, imported only by the tests that use it: importing it sets a default HIP
    queue count in os.environ, which is put back so later tests in this process see what they would
    have seen."""
    saved = os.environ.get("GPU_MAX_HW_QUEUES")
    sys.path.insert(0, ROOT)
    try:
        import bench as mod
    finally:
        sys.path.remove(ROOT)
        if saved is None:
            os.environ.pop("GPU_MAX_HW_QUEUES", None)
        else:
            os.environ["GPU_MAX_HW_QUEUES"] = saved
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line for a configuration that was not measured


def test_range_bytes_counts_every_leaf_walked(bench):
    # brute force: a range of length L reads its start and length (16 B), writes L (key, value) pairs
    # (16 B each) and walks the leaf holding its start plus one leaf per 8 keys it returns
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 300, size=257)
    want = 0
    for L in lens:
        leaves = 1 + -(-int(L) // 8)
        want += 16 + 16 * int(L) + 128 * leaves
    assert bench.range_bytes(lens) == want
    assert bench.range_bytes([0]) == 16 + 128
    assert bench.range_bytes([8]) == 16 + 128 + 256
    assert bench.range_bytes([9]) == 16 + 144 + 384


def test_byte_models_per_query(bench):
    # 8 B key in + 8 B value out + a 64-B key granule + a 64-B value granule of the leaf
    assert bench.BYTES_PER_QUERY == 8 + 8 + 64 + 64
    # + two radix passes, each reading and writing an 8-B key and a 4-B position
    assert bench.BYTES_PER_QUERY_PSA == bench.BYTES_PER_QUERY + 2 * 2 * (8 + 4)
