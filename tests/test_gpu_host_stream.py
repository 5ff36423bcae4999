"""Host-resident batches larger than one streaming chunk (2^22 queries): hm_search copies them through the
device chunk by chunk on two streams (copy in, sort, search, un-sort, copy out), overlapping one chunk's
copies with the other's kernels.  Every result is checked against the closed form (found exactly for keys
of the set, value = splitmix64_mix(key)) and against the device-buffer path, for pageable (numpy) and
pinned (torch pin_memory) caller memory, with ragged last chunks."""
import numpy as np
import pytest
import torch

from harmonia_amd import BPlusTree
from harmonia_amd import workloads as W

pytestmark = pytest.mark.gpu
CHUNK = 1 << 22


@pytest.fixture(scope="module")
def setup(gpu):
    keys, vals = W.gen_keys(1 << 22, 77)
    tree = BPlusTree.from_keys(keys, vals, fanout=64)
    sk = np.sort(keys)
    q = W.point_queries(sk, 2 * CHUNK + 123457, 78)  # three chunks, the last ragged; 90% hits
    found = np.isin(q, sk)
    exp_v = np.where(found, W.splitmix64_mix(q), np.uint64(0))
    yield tree, q, found.astype(np.uint8), exp_v
    tree.close()


def check(v, f, exp_f, exp_v):
    v = v.numpy().view(np.uint64) if hasattr(v, "numpy") else v
    f = f.numpy() if hasattr(f, "numpy") else f
    assert np.array_equal(f, exp_f)
    assert np.array_equal(v, exp_v)


@pytest.mark.parametrize("psa_bits", [-1, 0, 16])
def test_pageable_host_stream(setup, psa_bits):
    tree, q, exp_f, exp_v = setup
    v, f = tree.search_batch(q, psa_bits=psa_bits)
    check(v, f, exp_f, exp_v)


def test_pinned_host_stream(setup):
    tree, q, exp_f, exp_v = setup
    n = len(q)
    hq = torch.from_numpy(q.view(np.int64)).pin_memory()
    hv = torch.empty(n, dtype=torch.int64).pin_memory()
    hf = torch.empty(n, dtype=torch.uint8).pin_memory()
    hv.fill_(-1)
    hf.fill_(7)
    tree.search_batch(hq, out_vals=hv, out_found=hf)
    check(hv, hf, exp_f, exp_v)


def test_host_stream_equals_device_path(setup):
    tree, q, exp_f, exp_v = setup
    dq = torch.from_numpy(q.view(np.int64)).cuda()
    dv = torch.empty_like(dq)
    df = torch.empty(len(q), dtype=torch.uint8, device="cuda")
    tree.search_batch(dq, out_vals=dv, out_found=df)
    torch.cuda.synchronize()
    hv, hf = tree.search_batch(q)
    assert np.array_equal(df.cpu().numpy(), hf)
    assert np.array_equal(dv.cpu().numpy().view(np.uint64), hv)


@pytest.mark.parametrize("n", [CHUNK + 1, 2 * CHUNK, 4 * CHUNK + 64])
def test_host_stream_chunk_edges(setup, n):
    tree, q, exp_f, exp_v = setup
    reps = -(-n // len(q))
    qq = np.tile(q, reps)[:n]
    ef = np.tile(exp_f, reps)[:n]
    ev = np.tile(exp_v, reps)[:n]
    v, f = tree.search_batch(qq)
    check(v, f, ef, ev)
