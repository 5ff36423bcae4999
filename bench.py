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
sys.path.insert(0, ROOT)

METRIC = "Mqueries/s at 64M keys & % HBM roofline; served p50/p99 latency vs plain"
PEAK_HBM_GBS = 8000.0
BYTES_PER_QUERY = 144  # 8 B key in + 8 B value out + one 64-B leaf-key granule + one 64-B value granule
BYTES_PER_QUERY_PSA = 192  # + 2 radix passes over (8 B key + 4 B position), read and written


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


def cpu_baseline(tree, q_host, budget_s=12.0):
    """The oracle's restatement of the reference CPU search (CFB::search under an OpenMP
    parallel-for, H/regular_bpt/cfb.cpp:4-41, H/regular_bpt/main.cpp:92-116) on the same tree
    layout, over a bounded sample of the same query stream."""
    import oracle

    e = tree.export()
    s = e["stats"]
    lay = oracle.Layout(e["inner"], e["prefix"], e["leaves"], s.height, s.fanout,
                        np.array(s.level_start[:], dtype=np.uint32))
    cores = len(os.sched_getaffinity(0))
    sample = q_host[: 1 << 22]
    oracle.search_cfb(lay, sample[: 1 << 16], cores)  # warm-up
    times = []
    t_all = time.perf_counter()
    while time.perf_counter() - t_all < budget_s and len(times) < 5:
        t0 = time.perf_counter()
        v, f = oracle.search_cfb(lay, sample, cores)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(len(sample) / med / 1e6, 3), "unit": "Mqueries/s", "cores": cores, "kind": "port",
            "sample": f"{len(sample)} queries (first 2^22 of the same unsorted stream) on the same 2^26-key "
                      f"layout, median of {len(times)} runs, OpenMP static over queries"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--keys-log2", type=int, default=26)
    ap.add_argument("--queries-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--variants", default="all", help="'all' or 'headline'")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the tree is replicated per GPU "
              "and queries are sharded, so N GPUs is N independent copies of this run", file=sys.stderr)
        sys.exit(2)

    import torch

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

    variants = [("unsorted", 16), ("unsorted", 0), ("runs4096", 16), ("runs4096", 0)]
    if args.variants == "headline":
        variants = variants[:1]
    results = {}
    for order, psa in variants:
        fn = lambda o=order, p=psa: tree.search_batch(dq[o], out_vals=out_v, out_found=out_f, stream=stream,
                                                       psa_bits=p)
        ms, wall_ms = time_steps(fn, args.steps, args.warmup, stream)
        # kernel-level timing of the same call, HIP events on the launch stream
        lib.hm_profile_reset()
        lib.hm_profile_enable(1)
        for _ in range(max(3, args.steps // 4)):
            fn()
        torch.cuda.synchronize()
        search_ms, launches = stage_avg(lib, "search")
        hist_ms, _ = stage_avg(lib, "psa_hist")
        scat_ms, _ = stage_avg(lib, "psa_scatter")
        scan_ms, _ = stage_avg(lib, "scan")
        lib.hm_profile_enable(0)
        lib.hm_profile_reset()
        key = f"{order}_psa{psa}"
        results[key] = {"ms_per_step": round(ms, 4), "Mq_per_s": round(n_q / ms / 1e3, 1),
                        "search_kernel_ms": round(search_ms, 4) if search_ms else None,
                        "search_kernel_Mq_per_s": round(n_q / search_ms / 1e3, 1) if search_ms else None,
                        "psa_hist_ms": round(hist_ms, 4) if hist_ms else None,
                        "psa_scatter_ms": round(scat_ms, 4) if scat_ms else None,
                        "scan_ms": round(scan_ms, 4) if scan_ms else None}
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
                "kernel": "search_kernel (PSA-sorted queries, 2^24 per launch)",
                "bytes_per_query": BYTES_PER_QUERY}
    e2e_achieved = BYTES_PER_QUERY_PSA * n_q / (head["ms_per_step"] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(tree, q_unsorted)
        log("cpu", json.dumps(cpu))

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mqueries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "plain batched GPU search: fanout-64 tree of 2^26 uniform uint64 keys, 2^24 point "
                               "queries (90% hits), unsorted, PSA 16 bits", "keys": n_keys, "queries": n_q,
                   "fanout": 64, "leaf_keys": 8, "order": "unsorted", "psa_bits": 16, "hit_frac": 0.9},
        "roofline": roofline,
        "e2e_roofline_frac_psa192B": round(e2e_achieved / PEAK_HBM_GBS, 4),
        "cpu_baseline": cpu,
        "variants": results,
        "build_ms": round(build_ms, 1),
        "tree": {"height": st.height, "n_inner": st.n_inner, "n_leaves": st.n_leaves, "lds_levels": st.lds_levels},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
