#!/usr/bin/env python3
"""Batched point-search benchmark of the Harmonia B+tree on one MI355X.

Workload (default): the plain batched GPU search configuration -- fanout-64 tree bulk-built on
the device from 2^26 unique uint64 keys in [0, 2^63) with value = splitmix64_mix(key); 2^24 point
queries, 90% uniform hits and 10% misses, fully unsorted, PSA partial sort (top 16 bits) on.
The other orders / PSA settings of the same configuration are measured too and reported under
"variants".  A step = one hm_search call over the whole 2^24-query batch (sort + search + results
scattered back to input order), inputs and outputs resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# The served section runs several worker streams plus update epochs in this process: give the HIP
# runtime (which reads this when it initializes) a hardware queue per stream, so a read never queues
# behind a re-layout on a shared queue.  An explicit setting wins.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
sys.path.insert(0, ROOT)

METRIC = "Mqueries/s at 64M keys & % HBM roofline; served p50/p99 latency vs plain"
PEAK_HBM_GBS = 8000.0
BYTES_PER_QUERY = 144  # 8 B key in + 8 B value out + one 64-B leaf-key granule + one 64-B value granule
BYTES_PER_QUERY_PSA = 192  # + 2 radix passes over (8 B key + 4 B position), read and written
BYTES_PER_UPSERT = 272  # 16 B in + read and write of the leaf's key and value granules


def range_bytes(lens):
    """Algorithmic bytes of ranges of these lengths: start + length in, L (key, value) pairs out, and
    the key and value granules of every leaf walked (ceil(L/8) + 1 leaves)."""
    L = np.asarray(lens, dtype=np.int64)
    return int((16 + 16 * L + 128 * ((L + 7) // 8 + 1)).sum())


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps * 1e3


def stage_avg(lib, stage):
    import ctypes

    ms = ctypes.c_double(0)
    cnt = ctypes.c_uint64(0)
    lib.hm_profile_get(stage.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return (ms.value / cnt.value if cnt.value else None), int(cnt.value)


def cpu_model():
    """CPU model name as lscpu prints it (falls back to /proc/cpuinfo)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return None


def _median_time(fn, runs=5):
    """1 warm-up + `runs` timed calls; (median seconds, last result)."""
    r = fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        r = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), r


def cpu_leg(args):
    """The CPU reference path, run in a child process (OMP_PROC_BIND=spread set by the parent before
    the OpenMP runtime starts, so binding never touches the GPU-driving process).  Every result is
    compared with the sorted-array oracle before its time is reported.

    * the headline workload: the oracle's build of the same 2^26-key fanout-64 layout (bit-identical to
      the device build) and all 2^24 queries of the same unsorted stream, searched by the oracle's
      restatements of Harmonia's BPlusTree::search under search_cpu_test's OpenMP loop
      (H/src/main.cpp:772-812, H/src/bpt.cpp:673-680, H/src/avx.cpp:2-54) and of CFB::search under
      search_cpu_multithread's (H/regular_bpt/main.cpp:92-116, H/regular_bpt/cfb.cpp:4-41);
    * the CPU configuration (2^20 keys, fanout 64, 2^20 queries, 90% hits): the same two restatements
      plus, where oracle/_ref is built, the reference's own regular_bpt find and Harmonia
      BPlusTree::search compiled from its sources (trees built by its insert loop)."""
    import oracle
    from harmonia_amd import workloads as W

    cores = len(os.sched_getaffinity(0))
    res = {"unit": "Mqueries/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "timing": "1 warm-up + 5 timed runs, median"}

    def run_ports(lay, q, ev, ef):
        out = {}
        for name, fn in (("harmonia_search_port", lambda: oracle.search_harmonia(lay, q, cores)),
                         ("regular_find_port", lambda: oracle.search_cfb(lay, q, cores))):
            t, (v, f) = _median_time(fn)
            if not (np.array_equal(f, ef) and np.array_equal(v, ev)):
                raise SystemExit(f"cpu leg: {name} results differ from the oracle")
            out[name] = round(len(q) / t / 1e6, 3)
        return out

    # headline workload
    seed = W.config_seed(1)
    keys, vals = W.gen_keys(1 << args.keys_log2, seed)
    sk = np.sort(keys)
    del keys, vals
    sv = W.splitmix64_mix(sk)
    q = W.point_queries(sk, 1 << args.queries_log2, seed + 7, order="unsorted")
    lay = oracle.build_bfs(sk, sv, 64, 8)
    ev, ef = oracle.search_sorted(sk, sv, q)
    del sk, sv
    paths = run_ports(lay, q, ev, ef)
    del lay, q, ev, ef
    res["paths"] = paths
    res["value"] = max(paths.values())
    res["sample"] = (f"all {1 << args.queries_log2} queries of the same unsorted stream on the same "
                     f"2^{args.keys_log2}-key fanout-64 layout (built by the oracle from the same seed); value = "
                     "the faster of the two restated reference CPU searches, OpenMP static over queries")
    # CPU configuration: 2^20 keys, 2^20 queries
    seed0 = W.config_seed(0)
    keys, vals = W.gen_keys(1 << 20, seed0)
    sk = np.sort(keys)
    sv = W.splitmix64_mix(sk)
    q = W.point_queries(sk, 1 << 20, seed0 + 7)
    ev, ef = oracle.search_sorted(sk, sv, q)
    small = {"keys": 1 << 20, "queries": 1 << 20, "fanout": 64, "hit_frac": 0.9}
    small.update(run_ports(oracle.build_bfs(sk, sv, 64, 8), q, ev, ef))
    if oracle.ref_available():
        rt = oracle.RefRegularTree(keys, vals, 64)
        t, (v, f) = _median_time(lambda: rt.find(q, nthreads=cores))
        if not (np.array_equal(f, ef) and np.array_equal(v, ev)):
            raise SystemExit("cpu leg: reference regular_bpt find differs from the oracle")
        small["reference_regular_find"] = round(len(q) / t / 1e6, 3)
        ht = oracle.RefHarmoniaTree(keys, vals)
        t, hv = _median_time(lambda: ht.search(q, nthreads=cores))
        hit = ef.astype(bool)
        if not ((hv[~hit] == -1).all() and np.array_equal(hv[hit].view(np.uint64), ev[hit])):
            raise SystemExit("cpu leg: reference Harmonia search differs from the oracle")
        small["reference_harmonia_search"] = round(len(q) / t / 1e6, 3)
        small["reference_kind"] = "reference (its CPU sources compiled in oracle/_ref, trees built by its insert loop)"
    res["cpu_config"] = small
    print(json.dumps(res), flush=True)


def cpu_baseline(args):
    """Run cpu_leg in a child process with OMP_PROC_BIND=spread and one OpenMP thread per host core."""
    import subprocess

    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "spread"
    env["OMP_NUM_THREADS"] = str(len(os.sched_getaffinity(0)))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", "--keys-log2", str(args.keys_log2),
           "--queries-log2", str(args.queries_log2)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log(r.stderr[-2000:])
        raise SystemExit(f"cpu leg failed with status {r.returncode}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def gpu_cpu_config(dev, stream):
    """The CPU configuration (2^20 keys, fanout 64, 2^20 unsorted queries) on the GPU: one hm_search
    per step over device-resident queries, results in input order, checked against the oracle."""
    import torch

    import oracle
    from harmonia_amd import BPlusTree
    from harmonia_amd import workloads as W

    seed0 = W.config_seed(0)
    keys, vals = W.gen_keys(1 << 20, seed0)
    sk = np.sort(keys)
    q = W.point_queries(sk, 1 << 20, seed0 + 7)
    tree = BPlusTree.from_keys(keys, vals, fanout=64)
    dq = torch.from_numpy(q.view(np.int64)).to(dev)
    ov = torch.empty(len(q), dtype=torch.int64, device=dev)
    of = torch.empty(len(q), dtype=torch.uint8, device=dev)
    fn = lambda: tree.search_batch(dq, out_vals=ov, out_found=of, stream=stream)
    ms, _ = time_steps(fn, 20, 5, stream)
    fn()
    stream.synchronize()
    ev, ef = oracle.search_sorted(sk, W.splitmix64_mix(sk), q)
    ok = bool(np.array_equal(of.cpu().numpy(), ef) and np.array_equal(ov.cpu().numpy().view(np.uint64), ev))
    tree.close()
    if not ok:
        log("CPU-CONFIG GPU RESULT CHECK FAILED")
        sys.exit(1)
    return {"gpu_hm_search_Mq_per_s": round(len(q) / ms / 1e3, 1), "gpu_ms_per_call": round(ms, 4),
            "gpu_results_ok": ok}


def batched_uneven(dev, n_batches=10000, keys_log2=25, seed=4):
    """Uneven batches (sizes log-uniform in [32, 2^20], per-batch distribution uniform / Zipf(0.99) /
    sorted runs of 4096) on an autotuned 2^25-key tree: one hm_search per batch (the plain path) against
    the hipGraph-captured batcher that packs batches into one replay.  Both loops run in the C++ driver
    (harmonia_amd/libhm_batchbench.so); then every batch's results from a separate pass through the same
    batcher are checked against the closed form (found exactly for keys of the set, value =
    splitmix64_mix(key))."""
    import ctypes

    import torch

    from harmonia_amd import Batcher, BPlusTree
    from harmonia_amd import workloads as W

    cseed = W.config_seed(3) + seed
    keys, vals = W.gen_keys(1 << keys_log2, cseed)
    sk = np.sort(keys)
    pool_n = 1 << 24
    n_miss = pool_n - int(round(0.9 * pool_n))
    pools = [W.point_queries(sk, pool_n, cseed + 1),
             W.zipf_queries(sk, pool_n, cseed + 2),
             W.point_queries(sk, pool_n, cseed + 3, order="runs4096")]
    miss = [W.gen_misses(len(sk), n_miss, cseed + 1), None, W.gen_misses(len(sk), n_miss, cseed + 3)]
    g = W.rng(cseed, 8)
    sizes = W.log_uniform_sizes(n_batches, 32, 1 << 20, cseed).astype(np.uint64)
    dist = g.integers(0, 3, n_batches)
    starts = np.array([int(g.integers(0, pool_n - int(m) + 1)) for m in sizes], dtype=np.uint64)
    cap = 1 << 24
    # the tuner's sample is the first flush exactly as the batcher packs it (whole batches back to back,
    # each with its own order and skew), so it sees the batch size and locality a flush runs at
    fit = int(np.searchsorted(np.cumsum(sizes), cap, side="right"))
    sample = np.concatenate([pools[dist[j]][int(starts[j]):int(starts[j] + sizes[j])] for j in range(fit)])
    t0 = time.perf_counter()
    tree, tune = BPlusTree.from_keys_tuned(keys, vals, sample, quick=True)
    tune_s = time.perf_counter() - t0
    log("autotune:", json.dumps(tune), f"{tune_s:.1f} s")
    del sample, keys, vals
    dpools = [torch.from_numpy(p.view(np.int64)).to(dev) for p in pools]
    base = np.array([p.data_ptr() for p in dpools], dtype=np.uint64)
    ptrs = base[dist] + starts * np.uint64(8)
    total = int(sizes.sum())
    stream = torch.cuda.Stream()
    ov = torch.empty(1 << 20, dtype=torch.int64, device=dev)
    of = torch.empty(1 << 20, dtype=torch.uint8, device=dev)
    lib = ctypes.CDLL(os.path.join(ROOT, "harmonia_amd", "libhm_batchbench.so"))
    lib.hm_batch_bench_run.restype = ctypes.c_int
    lib.hm_batch_bench_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_uint64)]
    plain_s, batched_s, flushes = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_uint64(0)
    torch.cuda.synchronize()
    rc = lib.hm_batch_bench_run(tree._handle(), ptrs.ctypes.data, sizes.ctypes.data, n_batches, cap,
                                stream.cuda_stream, ov.data_ptr(), of.data_ptr(), ctypes.byref(plain_s),
                                ctypes.byref(batched_s), ctypes.byref(flushes))
    if rc:
        from harmonia_amd._lib import last_error
        raise SystemExit(f"hm_batch_bench_run failed: {rc} {last_error()}")
    # every batch checked, on a separate pass through a batcher of the same configuration
    exp = []
    for p, m in zip(pools, miss):
        f = np.ones(len(p), dtype=bool) if m is None else ~np.isin(p, m)
        exp.append((f.astype(np.uint8), np.where(f, W.splitmix64_mix(p), np.uint64(0))))
    b = Batcher(tree, cap, stream=stream)
    offs = np.zeros(n_batches, dtype=np.uint64)
    i = checked = mismatches = 0
    while i < n_batches:
        k = b.add_many(ptrs[i:], sizes[i:], offsets=offs[i:])
        b.flush()
        b.sync()
        fill = int(offs[i + k - 1] + sizes[i + k - 1])
        v_all, f_all = b.copy_out(0, fill)
        for j in range(i, i + k):
            m, d, s0, o = int(sizes[j]), int(dist[j]), int(starts[j]), int(offs[j])
            ok = (np.array_equal(f_all[o:o + m], exp[d][0][s0:s0 + m]) and
                  np.array_equal(v_all[o:o + m], exp[d][1][s0:s0 + m]))
            mismatches += 0 if ok else 1
            checked += 1
        i += k
    b.close()
    res = {"workload": f"{n_batches} batches, sizes log-uniform 32..2^20 (total {total} queries), per-batch "
                       f"uniform / Zipf(0.99) / runs of 4096, 2^{keys_log2}-key tree; both loops in C++",
           "plain_qps": round(total / plain_s.value, 1), "batched_qps": round(total / batched_s.value, 1),
           "speedup": round(plain_s.value / batched_s.value, 2), "plain_s": round(plain_s.value, 4),
           "batched_s": round(batched_s.value, 4), "flushes": int(flushes.value),
           "checked_batches": checked, "mismatches": mismatches, "autotune": tune, "autotune_s": round(tune_s, 1)}
    tree.close()
    if mismatches:
        log("UNEVEN-BATCH RESULT CHECK FAILED", mismatches)
        sys.exit(1)
    return res


def hbm_streaming(reps=10):
    """What plain streaming kernels reach on this GPU, beside the published 8 TB/s peak: torch's copy of a
    1-GiB int64 tensor (1 GiB read + 1 GiB written per launch) and its max-reduction over one (1 GiB read).
    HIP events on the current stream, where torch launches them; the average over `reps` launches after
    two warm-up launches."""
    import torch

    n = (1 << 30) // 8
    x = torch.arange(n, dtype=torch.int64, device="cuda:0")
    y = torch.empty_like(x)
    res = {}
    for name, fn, nbytes in (("copy", lambda: y.copy_(x), 2 * n * 8), ("read", lambda: x.max(), n * 8)):
        for _ in range(2):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b) / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "GB_per_s": round(gbs, 1), "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4)}
    del x, y
    torch.cuda.empty_cache()
    res["what"] = ("torch copy_ of 1 GiB (read + write) and max over 1 GiB (read), the practical streaming "
                   "rates that the search and range kernels' fractions of 8 TB/s compare with")
    return res


def host_streaming(tree, q, reps=3):
    """Queries and results in host memory (the reference's host-resident bucket loop,
    H/regular_bpt/prefix-sort.cu:261-309): hm_search streams a batch above 2^22 queries through the
    device in 2^22-query chunks on two streams, so the PCIe copies of one chunk overlap the other's
    kernels.  Pinned (torch pin_memory) and pageable (numpy) caller buffers; wall time of the whole call,
    median of `reps` after one warm-up; every result checked against the closed form."""
    import torch

    from harmonia_amd import workloads as W

    n = len(q)
    hq = torch.from_numpy(q.view(np.int64)).pin_memory()
    hv = torch.empty(n, dtype=torch.int64).pin_memory()
    hf = torch.empty(n, dtype=torch.uint8).pin_memory()
    pv = np.empty(n, dtype=np.uint64)
    pf = np.empty(n, dtype=np.uint8)
    runs = {"pinned": (lambda: tree.search_batch(hq, out_vals=hv, out_found=hf),
                       lambda: (hv.numpy().view(np.uint64), hf.numpy())),
            "pageable": (lambda: tree.search_batch(q, out_vals=pv, out_found=pf), lambda: (pv, pf))}
    res = {"workload": f"{n} queries and their results in host memory (2^26-key tree as above); chunks of "
                       "2^22 queries on two streams; PCIe bytes = 8 B in + 9 B out per query"}
    for name, (fn, get) in runs.items():
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        sec = float(np.median(ts))
        v, f = get()
        hit = f.astype(bool)
        ok = (int(hit.sum()) == round(0.9 * n) and np.array_equal(v[hit], W.splitmix64_mix(q[hit]))
              and not v[~hit].any())
        res[name] = {"ms": round(sec * 1e3, 3), "Mq_per_s": round(n / sec / 1e6, 1),
                     "pcie_GB_per_s": round(n * 17 / sec / 1e9, 2), "results_ok": bool(ok)}
    return res


def ablation(tree, dq, out_v, out_f, stream, steps, warmup):
    """The paper's ablation (Harmonia.pdf p.138-141) on this GPU, all on the same keys, leaves and
    2^24 unsorted queries: pointer-child trees (the reference's comparators), then the prefix-sum
    layout with the same kernel shape, then + PSA, + tuned narrowed groups, + LDS-resident levels.
    Every configuration's results are compared with the full pipeline's."""
    import torch

    n_q = dq.numel()
    tree.search_batch(dq, out_vals=out_v, out_found=out_f, stream=stream)
    stream.synchronize()
    ref_v, ref_f = out_v.cpu().numpy().copy(), out_f.cpu().numpy().copy()
    steps_cfg = [
        ("pointer_linear_g1", "regular tree, child pointers, one lane per query, linear node scan",
         lambda: tree.search_baseline(dq, out_v, out_f, kind="linear", group_size=1, stream=stream)),
        ("pointer_linear_g8", "noprefix-nosort: child pointers, 8-lane groups, linear node scan "
                              "(H/regular_bpt/noprefix-nosort.cu:89-237)",
         lambda: tree.search_baseline(dq, out_v, out_f, kind="linear", group_size=8, stream=stream)),
        ("pointer_index_g8", "HB+-tree: child pointers, 8-lane groups, index + key granule (H/src/gbpt.cu:34-113)",
         lambda: tree.search_baseline(dq, out_v, out_f, kind="pointer", group_size=8, stream=stream)),
        ("prefix_index_g8", "+ prefix-sum layout: the HB+ kernel with child = prefix[node] + slot",
         lambda: tree.search_baseline(dq, out_v, out_f, kind="prefix", group_size=8, stream=stream)),
        ("prefix_g8_persistent", "the production kernel at the same shape (8 lanes, 1 query per group, no LDS, "
                                 "no PSA): persistent XCD-mapped grid, 8 blocks per CU (one round at this shape)",
         lambda: tree.search_batch(dq, out_vals=out_v, out_found=out_f, stream=stream, group_size=8,
                                   queries_per_group=1, lds_levels=0, psa_bits=0, blocks_per_cu=8)),
        ("prefix_g8_psa16", "+ PSA partial sort (16 bits), results back in input order",
         lambda: tree.search_batch(dq, out_vals=out_v, out_found=out_f, stream=stream, group_size=8,
                                   queries_per_group=1, lds_levels=0, psa_bits=16, output_mode=1, blocks_per_cu=8)),
        ("prefix_psa16_ntg", "+ the tree's narrowed-group shape for PSA batches (lanes per query, queries in "
                             "flight per group)",
         lambda: tree.search_batch(dq, out_vals=out_v, out_found=out_f, stream=stream, lds_levels=0, psa_bits=16,
                                   output_mode=1)),
        ("full", "+ top levels staged in LDS (the default pipeline)",
         lambda: tree.search_batch(dq, out_vals=out_v, out_found=out_f, stream=stream, psa_bits=16, output_mode=1)),
    ]
    res = []
    for key, what, fn in steps_cfg:
        ms, _ = time_steps(fn, steps, warmup, stream)
        with torch.cuda.stream(stream):  # poison the flags on the stream the search runs on (ordered before it)
            out_f.fill_(7)
        fn()
        stream.synchronize()
        same = bool(np.array_equal(out_f.cpu().numpy(), ref_f) and np.array_equal(out_v.cpu().numpy(), ref_v))
        r = {"config": key, "what": what, "ms_per_step": round(ms, 4), "Mq_per_s": round(n_q / ms / 1e3, 1),
             "results_identical": same}
        log("ablation", json.dumps(r))
        res.append(r)
        if not same:
            log("ABLATION RESULT MISMATCH", key)
            sys.exit(1)
    return res


def skewed_and_ranges(dev, lib, steps, warmup):
    """Skewed search plus range scans: fanout-128 tree bulk-built from 2^26 keys; 2^24 point queries
    Zipf(0.99) over key rank (PSA 16, input order); 2^20 ranges, start uniform in [0, 2^63), length
    uniform in [16, 256] (range_prefix_sort_1thread_search_1_thread_scan_kernel, H/src/range.cu:101-227).
    Zipf results are checked in full (all hits, value = splitmix64_mix(key)); range counts in full
    against the closed form, pairs on a sample."""
    import torch

    from harmonia_amd import BPlusTree
    from harmonia_amd import workloads as W

    seed = W.config_seed(2)
    n_keys = 1 << 26
    keys, vals = W.gen_keys(n_keys, seed)
    tree = BPlusTree.from_keys(torch.from_numpy(keys.view(np.int64)).to(dev),
                               torch.from_numpy(vals.view(np.int64)).to(dev), fanout=128, leaf_fill=8)
    del vals
    sk = np.sort(keys)
    del keys
    stream = torch.cuda.Stream()
    q = W.zipf_queries(sk, 1 << 24, seed + 11)
    dq = torch.from_numpy(q.view(np.int64)).to(dev)
    ov = torch.empty(len(q), dtype=torch.int64, device=dev)
    of = torch.empty(len(q), dtype=torch.uint8, device=dev)
    fn = lambda: tree.search_batch(dq, out_vals=ov, out_found=of, stream=stream, psa_bits=16, output_mode=1)
    zms, _ = time_steps(fn, steps, warmup, stream)
    fn()
    stream.synchronize()
    ok = bool(of.all()) and np.array_equal(ov.cpu().numpy().view(np.uint64), W.splitmix64_mix(q))
    res = {"zipf_search": {"workload": "fanout-128 tree of 2^26 keys, 2^24 Zipf(0.99) point queries, PSA 16, "
                                       "results in input order", "ms_per_step": round(zms, 4),
                           "Mq_per_s": round(len(q) / zms / 1e3, 1), "results_ok": ok}}
    log("zipf", json.dumps(res["zipf_search"]))
    del dq, ov, of, q
    starts, lens = W.range_queries(1 << 20, seed + 13)
    n = len(starts)
    cap = int(lens.astype(np.int64).sum())
    d_s = torch.from_numpy(starts.view(np.int64)).to(dev)
    d_l = torch.from_numpy(lens.view(np.int32)).to(dev)
    d_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    d_k = torch.empty(cap, dtype=torch.int64, device=dev)
    d_v = torch.empty(cap, dtype=torch.int64, device=dev)
    d_c = torch.empty(n, dtype=torch.int32, device=dev)

    def rfn():
        rc = lib.hm_range(tree._handle(), d_s.data_ptr(), d_l.data_ptr(), n, d_off.data_ptr(), d_k.data_ptr(),
                          d_v.data_ptr(), d_c.data_ptr(), cap, stream.cuda_stream)
        if rc:
            raise RuntimeError(f"hm_range failed: {rc}")

    rms, _ = time_steps(rfn, steps, warmup, stream)
    lib.hm_profile_reset()
    lib.hm_profile_enable(1)
    for _ in range(3):
        rfn()
    torch.cuda.synchronize()
    kms, _ = stage_avg(lib, "range")
    lib.hm_profile_enable(0)
    lib.hm_profile_reset()
    lb = np.searchsorted(sk, starts, side="left").astype(np.int64)
    exp_cnt = np.minimum(lens.astype(np.int64), n_keys - lb)
    cnt = d_c.cpu().numpy().astype(np.int64)
    off = d_off.cpu().numpy()
    ok = bool(np.array_equal(cnt, exp_cnt))
    rk = d_k.cpu().numpy().view(np.uint64)
    rv = d_v.cpu().numpy().view(np.uint64)
    for i in W.rng(seed, 21).integers(0, n, 4096):
        c, o = int(exp_cnt[i]), int(off[i])
        ok &= bool(np.array_equal(rk[o:o + c], sk[lb[i]:lb[i] + c]))
        ok &= bool(np.array_equal(rv[o:o + c], W.splitmix64_mix(sk[lb[i]:lb[i] + c])))
    algo = range_bytes(exp_cnt)
    achieved = algo / (kms * 1e-3) / 1e9 if kms else None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "range_kernel_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    res["ranges"] = {"workload": "2^20 ranges, start uniform in [0, 2^63), length uniform in [16, 256], on the "
                                 "same tree", "ms_per_step": round(rms, 4),
                     "M_ranges_per_s": round(n / rms / 1e3, 2), "pairs_per_s_G": round(int(exp_cnt.sum()) / rms / 1e6, 2),
                     "range_kernel_ms": round(kms, 4) if kms else None,
                     "algorithmic_bytes": algo,
                     "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                                  "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                  "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
                                  "traffic": traffic, "kernel": "range_kernel<128> (2^20 ranges per launch)"},
                     "results_ok": ok}
    log("ranges", json.dumps(res["ranges"]))
    tree.close()
    if not (res["zipf_search"]["results_ok"] and ok):
        log("SKEWED / RANGE RESULT CHECK FAILED")
        sys.exit(1)
    return res


def batched_updates(dev, lib):
    """Batched upsert on the 2^26-key fanout-64 tree (bulk fill 6, so two free slots per leaf): the
    reference's update experiment shape, batches of 2^16 and 2^20 keys with 74.25% of them already in the
    tree (H/script/generate-data/generate.py:128-129; H/test.sh option 39), each applied as one epoch
    (hm_update + hm_epoch_publish, device buffers).  The last batch's keys are read back."""
    import torch

    from harmonia_amd import BPlusTree
    from harmonia_amd import workloads as W

    seed = W.config_seed(1)
    n_keys = 1 << 26
    keys, vals = W.gen_keys(n_keys, seed)
    tree = BPlusTree.from_keys(torch.from_numpy(keys.view(np.int64)).to(dev),
                               torch.from_numpy(vals.view(np.int64)).to(dev), fanout=64, leaf_fill=6)
    del vals
    cur = np.sort(keys)
    del keys
    stream = torch.cuda.Stream()
    g = W.rng(seed, 31)
    out = []
    step = 0
    # (batch size, share of existing keys, reported): the first two warm the scratch buffers up; each
    # reported configuration runs three timed epochs (fresh keys each; the median is reported), then
    # one more with the library's stage profiling on, untimed, for the re-layout time (the profiling
    # events would add their own latency to a timed epoch)
    plan = [(1 << 16, 0.7425, "warm-up"), (1 << 20, 0.7425, "warm-up")]
    for cfg in ((1 << 16, 1.0), (1 << 20, 1.0), (1 << 16, 0.7425), (1 << 20, 0.7425)):
        plan += [(cfg[0], cfg[1], "timed")] * 3 + [(cfg[0], cfg[1], "profiled")]
    prof_rl = []
    for batch, share, kind in plan:
        n_old = int(round(batch * share))
        old = cur[g.choice(len(cur), n_old, replace=False)]
        new = W.gen_misses(len(cur), max(batch - n_old, 1), seed + 100 + step)[: batch - n_old] ^ np.uint64(step << 56)
        new = np.unique(new)
        new = new[cur[np.minimum(np.searchsorted(cur, new), len(cur) - 1)] != new]
        bk = np.concatenate([old, new])
        g.shuffle(bk)
        bv = g.integers(0, 1 << 64, len(bk), dtype=np.uint64)
        dk = torch.from_numpy(bk.view(np.int64)).to(dev)
        dv = torch.from_numpy(bv.view(np.int64)).to(dev)
        torch.cuda.synchronize()
        prof = kind == "profiled"
        if prof:
            lib.hm_profile_reset()
            lib.hm_profile_enable(1)
        t0 = time.perf_counter()
        tree.update(dk, dv, stream=stream)
        rebuilt = tree.publish(stream=stream)
        stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        if prof:
            rl_ms, _ = stage_avg(lib, "relayout")
            lib.hm_profile_enable(0)
            lib.hm_profile_reset()
            prof_rl.append(round(rl_ms, 3) if rl_ms else None)
        cur = np.insert(cur, np.searchsorted(cur, new), new)  # new is sorted, unique and not in cur
        r = {"batch": len(bk), "existing": n_old, "new": len(new), "ms": round(ms, 3),
             "M_upserts_per_s": round(len(bk) / ms / 1e3, 2), "relaid_out": bool(rebuilt),
             "achieved_GBs_272B": round(BYTES_PER_UPSERT * len(bk) / (ms * 1e-3) / 1e9, 1)}
        log("update", json.dumps(r), "" if kind == "timed" else f"({kind})")
        if kind == "timed":
            out.append(r)
        step += 1
    batches = []
    for c, i in enumerate(range(0, len(out), 3)):  # median epoch of each configuration, with the spread
        reps = sorted(out[i:i + 3], key=lambda r: r["ms"])
        med = dict(reps[1])
        med["ms_min_max"] = [reps[0]["ms"], reps[-1]["ms"]]
        med["relaid_out_epochs"] = sum(1 for r in reps if r["relaid_out"])
        med["relayout_ms_profiled_epoch"] = prof_rl[c]
        batches.append(med)
    v, f = tree.search_batch(bk, psa_bits=0)
    ok = bool(f.all()) and np.array_equal(v, bv) and tree.stats().n_keys == len(cur)
    tree.close()
    if not ok:
        log("UPDATE RESULT CHECK FAILED")
        sys.exit(1)
    return {"workload": "2^26-key fanout-64 tree at fill 6; upsert batches of 2^16 / 2^20 keys, all existing "
                        "(values only, in place) or 74.25% existing, one epoch per batch, device buffers, after "
                        "one warm-up batch of each size; median of 3 epochs per configuration (timed without the "
                        "library's stage profiling; a fourth, profiled epoch gives the re-layout kernel time)",
            "batches": batches, "results_ok": ok}


def served(args):
    """Run the served-load driver (hm_serve_bench_main in harmonia_amd/libhm_servebench.so, called
    in-process): plain path closed-loop capacity, the plain path open-loop at 0.5 and 0.9 of it (its
    latency), then the serving loop open-loop at Poisson rates of 1, 3, 5 and 10 times that capacity."""
    import ctypes

    path = os.path.join(ROOT, "harmonia_amd", "libhm_servebench.so")
    if not os.path.exists(path):
        return {"error": "libhm_servebench.so not built"}
    lib = ctypes.CDLL(path)
    lib.hm_serve_bench_main.restype = ctypes.c_int
    lib.hm_serve_bench_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p,
                                        ctypes.c_uint64]
    common = ["--keys-log2", str(args.served_keys_log2), "--seed", "5"]

    def run(extra):
        argv = [b"hm_serve_bench"] + [a.encode() for a in common + extra]
        arr = (ctypes.c_char_p * len(argv))(*argv)
        out = ctypes.create_string_buffer(4096)
        rc = lib.hm_serve_bench_main(len(argv), arr, out, len(out))
        if not out.value:
            return {"error": f"rc={rc}"}
        res = json.loads(out.value.decode().strip())
        if rc != 0:
            res["error"] = f"rc={rc}"
        return res

    plain = run(["--plain", "--closed", "--requests", str(args.served_requests // 2)])
    log("plain", json.dumps(plain))
    res = {"plain_closed_loop": plain, "plain_open_loop": [], "served": []}
    if "error" in plain:
        return res
    base_rps = plain["requests_per_s"]
    # the plain path's own latency, open loop below its capacity (closed loop its p50 is queueing)
    for mult in (0.5, 0.9):
        r = run(["--plain", "--rate", str(base_rps * mult), "--requests", str(args.served_requests // 2)])
        r["rate_vs_plain_capacity"] = mult
        log("plain open loop", json.dumps(r))
        res["plain_open_loop"].append(r)
    for mult in (1.0, 3.0, 5.0, 10.0):
        r = run(["--rate", str(base_rps * mult), "--requests", str(args.served_requests)])
        r["rate_vs_plain_capacity"] = mult
        log("served", json.dumps(r))
        res["served"].append(r)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--keys-log2", type=int, default=26)
    ap.add_argument("--queries-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)  # the child process of cpu_baseline
    ap.add_argument("--variants", default="all",
                    help="'all', 'headline', or a comma-separated list of variant names (the headline always runs)")
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-batches comparison")
    ap.add_argument("--no-served", action="store_true", help="skip the served-load runs")
    ap.add_argument("--no-ablation", action="store_true", help="skip the layout / PSA / NTG ablation")
    ap.add_argument("--no-ranges", action="store_true", help="skip the fanout-128 Zipf search and range scans")
    ap.add_argument("--no-updates", action="store_true", help="skip the batched-update epochs")
    ap.add_argument("--no-host-stream", action="store_true", help="skip the host-memory streaming runs")
    ap.add_argument("--served-keys-log2", type=int, default=27)
    ap.add_argument("--served-requests", type=int, default=20000)
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a plain 1-GiB streaming kernel (read 1 GiB, write 1 GiB), the reference point "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the tree is replicated per GPU "
              "and queries are sharded, so N GPUs is N independent copies of this run", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate:
        n = (1 << 30) // 8
        x = torch.arange(n, dtype=torch.int64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.bitwise_xor(x, 1, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "bitwise_xor over 2^27 int64", "bytes_read": n * 8, "bytes_written": n * 8,
                          "launches": 3}))
        return

    import harmonia_amd
    from harmonia_amd import BPlusTree
    from harmonia_amd import workloads as W

    lib = harmonia_amd.load()
    dev = torch.device("cuda:0")
    n_keys, n_q = 1 << args.keys_log2, 1 << args.queries_log2
    seed = W.config_seed(1)
    log(f"generating {n_keys} keys / {n_q} queries")
    keys, vals = W.gen_keys(n_keys, seed)
    sk = np.sort(keys)
    q_unsorted = W.point_queries(sk, n_q, seed + 7, order="unsorted")
    q_runs = W.order_queries(q_unsorted, "runs4096", seed + 7)
    del sk
    dk = torch.from_numpy(keys.view(np.int64)).to(dev)
    dvals = torch.from_numpy(vals.view(np.int64)).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = BPlusTree.from_keys(dk, dvals, fanout=64, leaf_fill=8)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    del dk, dvals
    st = tree.stats()
    log(f"built: height {st.height} inner {st.n_inner} leaves {st.n_leaves} lds_levels {st.lds_levels} "
        f"in {build_ms:.1f} ms")

    stream = torch.cuda.Stream()
    out_v = torch.empty(n_q, dtype=torch.int64, device=dev)
    out_f = torch.empty(n_q, dtype=torch.uint8, device=dev)
    dq = {"unsorted": torch.from_numpy(q_unsorted.view(np.int64)).to(dev),
          "runs4096": torch.from_numpy(q_runs.view(np.int64)).to(dev)}
    torch.cuda.synchronize()

    # (order, PSA bits, output mode): mode 1 = results in input order (the headline), 3 = PSA order plus
    # positions (the reference's own contract)
    # sorted kernel (PSA batches): 0 = the tree's choice (one descent per query), 2 = the leaf window
    variants = [("unsorted", 16, 1, 0), ("unsorted", 16, 1, 2), ("unsorted", 0, 1, 0),
                ("runs4096", 16, 1, 0), ("runs4096", 0, 1, 0), ("unsorted", 16, 3, 0)]
    def vkey(order, psa, mode, skern):
        return (f"{order}_psa{psa}" + ("_psa_order" if mode == 3 else "") +
                {0: "", 1: "_per_query_kernel", 2: "_leaf_window_kernel"}[skern])

    if args.variants == "headline":
        variants = variants[:2]
    elif args.variants != "all":  # a comma-separated list of variant names
        want = set(args.variants.split(","))
        variants = [v for v in variants if vkey(*v) in want or v == variants[0]]
    out_pos = torch.empty(n_q, dtype=torch.int32, device=dev)
    results = {}
    for order, psa, mode, skern in variants:
        fn = lambda o=order, p=psa, m=mode, k=skern: tree.search_batch(
            dq[o], out_vals=out_v, out_found=out_f, stream=stream, psa_bits=p, output_mode=m if p else 0,
            out_pos=out_pos if m == 3 else None, sorted_kernel=k)
        ms, wall_ms = time_steps(fn, args.steps, args.warmup, stream)
        # kernel-level timing of the same call, HIP events on the launch stream
        lib.hm_profile_reset()
        lib.hm_profile_enable(1)
        reps = max(3, args.steps // 4)
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        search_ms, launches = stage_avg(lib, "search")  # one search launch per step

        def per_step(stage):
            avg, cnt = stage_avg(lib, stage)
            return avg * cnt / reps if avg else None

        hist_ms, scat_ms, scan_ms, unsort_ms = (per_step(x) for x in ("psa_hist", "psa_scatter", "scan", "unsort"))
        lib.hm_profile_enable(0)
        lib.hm_profile_reset()
        key = vkey(order, psa, mode, skern)
        results[key] = {"ms_per_step": round(ms, 4), "Mq_per_s": round(n_q / ms / 1e3, 1),
                        "search_kernel_ms": round(search_ms, 4) if search_ms else None,
                        "search_kernel_Mq_per_s": round(n_q / search_ms / 1e3, 1) if search_ms else None,
                        "psa_hist_ms": round(hist_ms, 4) if hist_ms else None,
                        "psa_scatter_ms": round(scat_ms, 4) if scat_ms else None,
                        "scan_ms": round(scan_ms, 4) if scan_ms else None,
                        "unsort_ms": round(unsort_ms, 4) if unsort_ms else None}
        log(key, json.dumps(results[key]))

    # correctness spot check of the headline output (size-independent property)
    tree.search_batch(dq["unsorted"], out_vals=out_v, out_found=out_f, stream=stream, psa_bits=16)
    stream.synchronize()
    f = out_f.cpu().numpy()
    v = out_v.cpu().numpy().view(np.uint64)
    hit = f.astype(bool)
    ok = int(hit.sum()) == round(0.9 * n_q) and np.array_equal(v[hit], W.splitmix64_mix(q_unsorted[hit]))
    if not ok:
        log("RESULT CHECK FAILED")
        sys.exit(1)

    head = results["unsorted_psa16"]
    value = n_q / head["ms_per_step"] / 1e3  # Mqueries/s
    sk_ms = head["search_kernel_ms"]
    achieved = BYTES_PER_QUERY * n_q / (sk_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "search_kernel_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": "search_span_kernel<64,OUT_DIRECT> (PSA-sorted queries, 2^24 per launch)",
                "bytes_per_query": BYTES_PER_QUERY}
    e2e_achieved = BYTES_PER_QUERY_PSA * n_q / (head["ms_per_step"] * 1e-3) / 1e9
    # the whole step (sort + search + un-sort, results in input order) against the 192-B/query model
    roofline_e2e = {"bound": "hbm", "achieved": round(e2e_achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(e2e_achieved / PEAK_HBM_GBS, 4), "bytes_per_query": BYTES_PER_QUERY_PSA,
                    "what": "whole hm_search step: PSA sort + search + un-sort to input order"}

    hbm = hbm_streaming()
    log("hbm_streaming", json.dumps(hbm))

    hstream = None
    if not args.no_host_stream:
        log("host streaming ...")
        hstream = host_streaming(tree, q_unsorted)
        log("host_streaming", json.dumps(hstream))
        if not (hstream["pinned"]["results_ok"] and hstream["pageable"]["results_ok"]):
            log("HOST STREAMING RESULT CHECK FAILED")
            sys.exit(1)

    abl = None
    if not args.no_ablation:
        log("ablation ...")
        abl = ablation(tree, dq["unsorted"], out_v, out_f, stream, max(3, args.steps // 2), 2)

    tree.close()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(args)
        cpu["cpu_config"].update(gpu_cpu_config(dev, stream))
        log("cpu", json.dumps(cpu))
    del dq, out_v, out_f
    torch.cuda.empty_cache()

    skewed = None
    if not args.no_ranges:
        log("skewed search + ranges ...")
        skewed = skewed_and_ranges(dev, lib, max(3, args.steps // 2), 2)
        torch.cuda.empty_cache()
    updates = None
    if not args.no_updates:
        log("batched updates ...")
        updates = batched_updates(dev, lib)
        torch.cuda.empty_cache()

    batched = None
    if not args.no_batched:
        log("uneven batches ...")
        batched = batched_uneven(dev)
        log("batched", json.dumps({k: v for k, v in batched.items() if k != "autotune"}))
        torch.cuda.empty_cache()
    serve = None
    if not args.no_served:
        log("served load ...")
        serve = served(args)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mqueries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "plain batched GPU search: fanout-64 tree of 2^26 uniform uint64 keys, 2^24 point "
                               "queries (90% hits), unsorted, PSA 16 bits, results in input order", "keys": n_keys,
                   "queries": n_q,
                   "fanout": 64, "leaf_keys": 8, "order": "unsorted", "psa_bits": 16, "hit_frac": 0.9},
        "roofline": roofline,
        "roofline_e2e": roofline_e2e,
        "cpu_baseline": cpu,
        "variants": results,
        "build_ms": round(build_ms, 1),
        "tree": {"height": st.height, "n_inner": st.n_inner, "n_leaves": st.n_leaves, "lds_levels": st.lds_levels},
        "hbm_streaming": hbm,
        "host_streaming": hstream,
        "ablation": abl,
        "skewed_and_ranges": skewed,
        "updates": updates,
        "batched_uneven": batched,
        "served_mixed": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
