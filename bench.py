#!/usr/bin/env python3
"""QSketch on one MI355X: insert + query throughput at Zipf alpha = 1.1.

Workload (one "step"): clear the sketch, insert a stream of 2^28 uint32 keys
drawn Zipf(1.1) over a 2^26-key universe (rank -> key by a seeded permutation),
then query every distinct key of the stream.  The sketch uses the reference
sizing rule at work-load factor 0.5 (n_low = 201,326,557 buckets = 6 GiB) and a
2^18-bucket high pool.  Keys are generated on the GPU before the timed region.

Prints one JSON line.  value = (inserted + queried keys) / wall time of the K
timed steps, in G keys/s.  The roofline object is for the optimistic insert
kernel (the dominant one): algorithmic bytes 68 B per key (4 B key + 32 B bucket
read + 32 B bucket write), divided by its summed launch durations measured with
HIP events on its launch stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
INSERT_BYTES_PER_KEY = 68   # u32 key + 32 B bucket read + 32 B bucket write
QUERY_BYTES_PER_KEY = 40    # u32 key + 32 B bucket read + 4 B estimate
# HBM bytes per key the insert path moved on the default workload, from the
# FETCH_SIZE / WRITE_SIZE counter runs calibrated with a 1 GiB copy
# (profiles/insert_query_plain_zipf1.1_2e28_traffic.md)
INSERT_TRAFFIC_BYTES_PER_KEY = 49.5
DEFAULT_WORKLOAD = dict(log2n=28, log2u=26, alpha=1.1, wlf=0.5, n_high=1 << 18)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=28, help="log2 of the stream length")
    ap.add_argument("--log2u", type=int, default=26, help="log2 of the key universe")
    ap.add_argument("--alpha", type=float, default=1.1, help="Zipf exponent; 0 = uniform keys")
    ap.add_argument("--wlf", type=float, default=0.5, help="work-load factor of the sizing rule")
    ap.add_argument("--n-high", type=int, default=1 << 18)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the learned-form comparison (ARE at equal memory) of the default line")
    ap.add_argument("--cpu-sample-log2", type=int, default=27,
                    help="log2 of the stream prefix the CPU baseline inserts (2^27: about 20 s of "
                         "thread time on 16 host threads)")
    ap.add_argument("--workload", choices=["plain", "multi", "learned", "sf"], default="plain",
                    help="plain: one sketch (default); multi: 4096 independent sketches in one batch; "
                         "learned: MLP-routed heavy hitters + sketch vs plain sketch at equal memory; "
                         "sf: the SF sketch with split counters (m = 8, 3 splits)")
    ap.add_argument("--mem-scales", type=str, default="1,0.25,0.0625",
                    help="learned: sketch memory budgets as fractions of the sizing rule's")
    ap.add_argument("--train-steps", type=int, default=3000, help="learned: Adam steps of 4096 pairs")
    ap.add_argument("--sketches", type=int, default=4096, help="multi: number of sketches")
    ap.add_argument("--no-tune", action="store_true", help="skip the segment-pass shape autotuner")
    ap.add_argument("--no-hot-filter", action="store_true",
                    help="insert every key through the ordered pipeline (same result, for comparison)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first copy 1 GiB device to device (a known-traffic kernel for counter calibration)")
    ap.add_argument("--calib", action="store_true",
                    help="only measure the ceilings: a streaming copy and random reads of 16/32/64/128 B "
                         "from a 6 GiB buffer (include/qsketch_calib.h); one JSON line")
    return ap.parse_args()


def cpu_info() -> dict:
    """CPU model and clock of this host, read at run time from /proc/cpuinfo."""
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model is None:
                    model = v.strip()
                elif k == "cpu MHz":
                    mhz.append(float(v))
    except OSError:
        pass
    return {"model": model, "mhz_mean": round(sum(mhz) / len(mhz), 1) if mhz else None,
            "logical_cpus": os.cpu_count()}


def host_threads():
    """Threads for the all-cores variant: every core of the affinity mask, but
    no more than OMP_NUM_THREADS when that is set (a host shared between
    GPUs may give one GPU's job a share of its cores and say so there)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else aff
    return max(1, min(aff, cap)), aff


def _oracle_times(keys, uniq, n_low, n_high, threads, reps):
    """Median insert and query seconds of the oracle over `reps` runs, each on
    a cleared sketch whose pages are already touched."""
    import numpy as np
    from oracle import pyoracle as po
    o = po.OracleSketch(n_low, n_high, 1021, rng_seed=1)
    ti, tq = [], []
    for _ in range(reps):
        o.clear()
        t0 = time.perf_counter()
        rc = o.insert(keys, nthreads=threads)
        t1 = time.perf_counter()
        o.query(uniq, nthreads=threads)
        t2 = time.perf_counter()
        if rc != 0:
            raise RuntimeError(f"oracle insert failed: {rc}")
        ti.append(t1 - t0)
        tq.append(t2 - t1)
    return float(np.median(ti)), float(np.median(tq))


def _rate(n, nq, ti, tq) -> dict:
    return {"gkeys_s": round((n + nq) / (ti + tq) / 1e9, 5), "insert_mkeys_s": round(n / ti / 1e6, 2),
            "query_mkeys_s": round(nq / tq / 1e6, 2), "insert_s": round(ti, 4), "query_s": round(tq, 4)}


def cpu_baseline(keys_dev, n_low, n_high, sample_log2):
    """The oracle (the restated sequential semantics; the reference has no CPU
    sketch) timed on this host's cores:
      * the measured stream: its first 2^sample_log2 keys inserted, then their
        distinct keys queried, all threads (bucket-range partitioned, the same
        state as one thread), median of 3;
      * cfg1 (2^22 keys, Zipf 1.0 over 2^18, seed 1, WLF 0.5), one thread and
        all threads, median of 5 each."""
    import numpy as np
    import qsketch_amd as q
    from qsketch_amd import workloads as wl
    threads, aff = host_threads()
    n = min(1 << sample_log2, keys_dev.numel())
    keys = keys_dev[:n].cpu().numpy().view(np.uint32).copy()
    uniq = np.unique(keys)
    ti, tq = _oracle_times(keys, uniq, n_low, n_high, threads, 3)
    main_rate = _rate(n, uniq.size, ti, tq)
    k1 = wl.plain_keys(**wl.CFG1, device=False)
    u1 = np.unique(k1)
    n_low1 = q.number_of_buckets(k1.size, 8, 0.5)
    one = _rate(k1.size, u1.size, *_oracle_times(k1, u1, n_low1, 1 << 13, 1, 5))
    allc = _rate(k1.size, u1.size, *_oracle_times(k1, u1, n_low1, 1 << 13, threads, 5))
    return {
        "value": main_rate["gkeys_s"],
        "unit": "Gkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first 2^{sample_log2} keys of the same stream inserted, then their {uniq.size} distinct "
                  f"keys queried; bucket-range partitioned oracle on {threads} threads, median of 3 "
                  f"(insert {main_rate['insert_mkeys_s']} Mkeys/s)",
        "threads_note": f"all-cores variant on {threads} threads: affinity mask {aff} cores, "
                        f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}",
        "cpu": cpu_info(),
        "measured_stream": main_rate,
        "cfg1": {"workload": "2^22 u32 keys, Zipf 1.0 over 2^18, seed 1; n_low 3,145,721 (WLF 0.5), "
                             "n_high 2^13; insert all, query every distinct key; median of 5",
                 "distinct": int(u1.size), "single_core": one, "all_cores": dict(allc, threads=threads)},
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; sharding across GPUs is "
                                   "designed but not built", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.calib_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if a.calib:
        return bench_calib(a)
    if a.workload == "multi":
        return bench_multi(a)
    if a.workload == "learned":
        return bench_learned(a)
    if a.workload == "sf":
        return bench_sf(a)
    import torch
    import qsketch_amd as q

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = 1 << a.log2n
    U = 1 << a.log2u
    if a.alpha > 0:
        z = q.Zipf(a.alpha, U)
        keys = z.keys_u32(seed=1, n=N)
        dist = f"zipf{a.alpha}"
    else:
        keys = q.uniform_u32(seed=1, n=N)
        dist = "uniform"
    uniq = torch.unique(keys)
    counts = torch.empty_like(uniq)
    n_low = q.number_of_buckets(N, 8, a.wlf)
    sk = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, a.n_high, 32, 1, 1021, rng_seed=1)
    sk.set_hot_filter(not a.no_hot_filter)  # the tuner times the insert as it will run
    tune = None if a.no_tune else sk.autotune(keys, reps=1)
    torch.cuda.synchronize()

    def step():
        sk.clear()
        sk.insert(keys, N)
        sk.search(uniq, uniq.numel(), counts)

    for _ in range(a.warmup):
        step()
    sk.set_timing(True)
    torch.cuda.synchronize()
    step_s = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        step()
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tim = sk.timing()
    st = sk.stats()
    sk.set_timing(False)

    dt = t1 - t0
    n_q = uniq.numel()
    value = a.steps * (N + n_q) / dt / 1e9
    ins_keys = tim["insert_keys"]
    ins_ms = tim["insert_kernel_ms"]
    achieved = ins_keys * INSERT_BYTES_PER_KEY / (ins_ms * 1e-3) / 1e9 if ins_ms > 0 else 0.0
    q_ms = tim["query_kernel_ms"]
    q_achieved = tim["query_keys"] * QUERY_BYTES_PER_KEY / (q_ms * 1e-3) / 1e9 if q_ms > 0 else 0.0
    kern = {k: v for k, v in tim["kernels"].items() if k != "query" and v["launches"]}
    dom = max(kern, key=lambda k: kern[k]["ms"]) if kern else None
    # one "launch" of the insert path = the kernel sequence of one insert call
    # of N keys (the hot-key filter, then the ordered pipeline on the prefix
    # and on the keys the filter leaves)
    keys_per_launch = N
    path_ms = ins_ms / a.steps
    default = all(getattr(a, k) == v for k, v in DEFAULT_WORKLOAD.items())
    traffic = round(INSERT_TRAFFIC_BYTES_PER_KEY * keys_per_launch) if default else None
    # the radix scatter on its own: it moves one 8-byte item in and out per
    # sorted key and pass (the first pass reads the 4-byte key), and every pass
    # but the last writes the next pass's one-byte digit; the filter's hot
    # keys are never sorted
    sd = kern.get("sort_down")
    dom_obj = None
    hot_per_step = st["hot_keys"]  # counted since the last clear: the last step's insert
    sorted_items = N - hot_per_step
    if sd:
        chunks = max(1, kern.get("seg", {"launches": 1})["launches"])
        passes = sd["launches"] / chunks
        item_bytes = 4 + 8 + (passes - 1) * 16 + (passes - 1)
        sd_ms = sd["ms"] / a.steps
        dom_obj = {"kernel": "k_bs_down (radix scatter)", "ms_per_step": round(sd_ms, 4),
                   "launches_per_step": sd["launches"] / a.steps, "sorted_keys_per_step": round(sorted_items),
                   "algorithmic_bytes_per_step": round(item_bytes * sorted_items),
                   "achieved": round(item_bytes * sorted_items / (sd_ms * 1e-3) / 1e9, 1),
                   "frac": round(item_bytes * sorted_items / (sd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out = {
        "metric": "insert & query Gkeys/s (% HBM roofline) at Zipf 1.1; ARE at equal memory",
        "value": round(value, 4),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain QSketch insert 2^{a.log2n} u32 keys {dist} over 2^{a.log2u} keys, "
                        f"then query all {n_q} distinct keys",
            "n_low": n_low, "n_high": a.n_high, "wlf": a.wlf, "keys": N, "distinct": n_q,
        },
        "step_ms": [round(x * 1e3, 3) for x in step_s],
        "hot_filter": not a.no_hot_filter,
        "insert_gkeys_s": round(ins_keys / (ins_ms * 1e-3) / 1e9, 3) if ins_ms > 0 else None,
        "query_gkeys_s": round(tim["query_keys"] / (q_ms * 1e-3) / 1e9, 3) if q_ms > 0 else None,
        "roofline": {
            "bound": "hbm",
            "kernel": "insert path: on the prefix and on the cold keys, per radix pass k_bs_up (first "
                      "pass) or k_bs_up_dig, k_hist_scan_local, k_hist_scan_groups, k_bs_down; then k_segw, "
                      "k_ordered, k_trig_keys + pair sort, k_assign, k_tail, k_clear_long, k_next, k_reset; "
                      "between them the hot-key filter (k_hot_build, k_hot_count, k_scan_*, k_hot_move, "
                      "k_hot_flush).  The filter's candidate sample sort runs on a second stream beside the "
                      "prefix; any wait for it is counted in the filter's time",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "not measured in this run: bytes per key from the FETCH_SIZE/WRITE_SIZE "
                              "counter runs of this workload (calibrated by a 1 GiB copy) in "
                              "profiles/insert_query_plain_zipf1.1_2e28_traffic.md, times the keys "
                              "of one launch" if traffic else None,
            "launch": "the kernel sequence of one insert call of all N keys (hot-key filter; radix "
                      "passes, k_segw, k_ordered, k_assign, k_tail, ... on the prefix and on the keys "
                      "the filter leaves), timed with HIP events around each kernel on its stream",
            "bytes_per_key": INSERT_BYTES_PER_KEY,
            "keys_per_launch": round(keys_per_launch),
            "avg_launch_ms": round(path_ms, 4),
            "path_ms_per_step": round(path_ms, 4),
            "hot_filtered_keys_per_step": round(hot_per_step),
            "dominant_kernel": dom,
            "dominant": dom_obj,
            "kernel_ms_per_step": {k: round(v["ms"] / a.steps, 4) for k, v in kern.items()},
        },
        "roofline_query": {
            "bound": "hbm", "kernel": "k_query<4>", "achieved": round(q_achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(q_achieved / HBM_PEAK_GBS, 4),
            "bytes_per_key": QUERY_BYTES_PER_KEY,
            "avg_launch_ms": round(q_ms / max(1, tim["query_kernel_launches"]), 4),
        },
        "shape": list(sk.shape()),
        "tuner_best": tune["best"] if tune else None,
        "fixup": {"keys_replayed_per_step": st["fixup_keys"],
                  "segments_replayed_per_step": st["ordered_segments"],
                  "long_segments_per_step": st["long_segments"],
                  "linked_buckets": st["linked_buckets"],
                  "pool_occupancy": round(st["linked_buckets"] / (a.n_high - 1), 4),
                  "allocator": "ids assigned in stream order by a sort of the triggering inserts; "
                               "no compare-and-swap loop, so no retries"},
    }
    # the measured ceilings of the two dominant access patterns, live
    ceil = ceilings(copy_gib=(4,), gathers=(32,))
    copy_gbs = ceil["copy"]["4GiB"]["gb_s"]
    g32 = ceil["gather"]["32B"]["greads_s"]
    out["roofline_query"].update({
        "ceiling_random_32B_greads_s": g32,
        "frac_of_random_32B_ceiling": round(out["query_gkeys_s"] / g32, 4) if out["query_gkeys_s"] else None,
        "ceiling_note": "2^25 random 32-byte reads from a 6 GiB buffer, 4-byte result each "
                        "(qs_calib_gather), measured in this run; every random read moves a whole "
                        "128-byte line (profiles/calibration_copy_fill_gather.md)"})
    if out["roofline"]["dominant"]:
        out["roofline"]["dominant"]["frac_of_copy_ceiling"] = round(out["roofline"]["dominant"]["achieved"] / copy_gbs, 4)
    out["roofline"]["copy_ceiling_gb_s"] = copy_gbs
    out["roofline"]["frac_of_copy_ceiling"] = round(out["roofline"]["achieved"] / copy_gbs, 4)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(keys, n_low, a.n_high, a.cpu_sample_log2)
    if not a.no_accuracy:
        sk.destroy()
        del keys, uniq, counts, sk
        torch.cuda.empty_cache()
        out["accuracy"] = learned_accuracy(a)
    print(json.dumps(out), flush=True)


def bench_multi(a):
    """Batched multi-sketch: one captured graph launch inserts every stream into
    its own sketch, against the same streams inserted by one plain sketch each."""
    import numpy as np
    import torch
    import qsketch_amd as q

    from qsketch_amd import workloads as wl

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    K = a.sketches
    keys, offs, lengths, alphas, seeds = wl.multi_keys(K)
    N = int(offs[-1])
    d_offs = torch.from_numpy(offs.view(np.int64).copy()).to(dev)
    params = wl.multi_params(lengths, seeds)
    S = q.SketchSet(params)
    tune = None
    if not a.no_tune:
        tune = S.autotune(keys, offs, d_offs, reps=1)
    g = S.capture_insert(keys, offs, d_offs)
    torch.cuda.synchronize()

    def run_set():
        S.clear()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.launch(0)
        S.sync()
        return time.perf_counter() - t0

    for _ in range(a.warmup):
        run_set()
    t_set = [run_set() for _ in range(a.steps)]
    # per-kernel breakdown: the same launch sequence, uncaptured, with events
    S.clear()
    S.set_timing(True)
    S.insert_async(keys, offs, d_offs, 0)
    S.sync()
    tim = S.timing()
    S.set_timing(False)

    # the same streams, one plain sketch and one blocking insert call each
    plain = [q.Sketch_GPU_Sub_Warp_Level(int(p.n_low), 8, int(p.n_high), 32, rng_seed=int(p.rng_seed))
             for p in params]
    views = [keys[int(offs[i]):int(offs[i + 1])] for i in range(K)]

    def run_seq(blocking):
        for sk in plain:
            sk.clear()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for sk, v in zip(plain, views):
            if blocking:
                sk.insert(v)
            else:
                sk.insert_async(v, v.numel(), 0)
        torch.cuda.synchronize()
        for sk in plain:
            sk.sync()
        return time.perf_counter() - t0

    run_seq(True)
    t_seq = min(run_seq(True) for _ in range(max(1, a.steps)))
    t_seq_async = min(run_seq(False) for _ in range(max(1, a.steps)))

    # parity: every sketch of the set gives the plain sketch's estimates on all
    # distinct keys of its stream
    uq = [torch.unique(v) for v in views]
    qoffs = np.zeros(K + 1, np.uint64)
    qoffs[1:] = np.cumsum([u.numel() for u in uq])
    qkeys = torch.cat(uq)
    est_set = S.search(qkeys, qoffs)
    est_plain = torch.cat([sk.search(u) for sk, u in zip(plain, uq)])
    identical = bool(torch.equal(est_set, est_plain))
    st = S.stats()

    dt = float(np.median(t_set))
    kern_ms = sum(v["ms"] for k, v in tim["kernels"].items() if k != "query" and v["launches"])
    out = {
        "metric": "batched multi-sketch insert Gkeys/s (one captured launch) vs sequential plain inserts",
        "value": round(N / dt / 1e9, 4),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"{K} sketches, stream lengths log-uniform in [2^10, 2^20] u32 keys, Zipf alpha "
                        f"uniform in [0.7, 1.3] (0.01 grid), own seed and 2^16-key universe each; "
                        f"n_low per sketch by the sizing rule at WLF 0.5, n_high max(1024, n/256)",
            "keys": N, "sketches": K, "memory_bytes": S.memory_bytes(),
        },
        "sequential_blocking": {"gkeys_s": round(N / t_seq / 1e9, 4), "ms": round(t_seq * 1e3, 3),
                                "calls": K},
        "sequential_async": {"gkeys_s": round(N / t_seq_async / 1e9, 4), "ms": round(t_seq_async * 1e3, 3)},
        "speedup_vs_sequential_blocking": round(t_seq / dt, 2),
        "estimates_identical_to_plain_sketches": identical,
        "distinct_keys_checked": int(qoffs[-1]),
        "shape": list(S.shape()),
        "tuner": tune,
        "kernel_ms_per_step": {k: round(v["ms"], 4) for k, v in tim["kernels"].items() if v["launches"]},
        "kernel_ms_sum": round(kern_ms, 4),
        "launch_overhead_ms": round(dt * 1e3 - kern_ms, 4),
        "linked_buckets": st["linked_buckets"],
        "roofline": {
            "bound": "hbm", "kernel": "the captured insert launch sequence of the set (radix passes, "
                                      "k_segw, k_ordered, k_assign, k_tail, ...)",
            "bytes_per_key": INSERT_BYTES_PER_KEY,
            "achieved": round(N * INSERT_BYTES_PER_KEY / dt / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(N * INSERT_BYTES_PER_KEY / dt / 1e9 / HBM_PEAK_GBS, 4),
            "avg_launch_ms": round(dt * 1e3, 4), "traffic": None,
        },
    }
    g.destroy()
    print(json.dumps(out), flush=True)


def ceilings(reps=5, copy_gib=(1, 4), gathers=(16, 32, 64, 128)):
    """Ceilings of the access patterns the sketch kernels are made of, with
    kernels whose traffic is known (qsketch_amd/csrc/qs_calib.hip), timed
    with events on the stream they run on:
      * a streaming copy (read + write) of 1 GiB and of 4 GiB;
      * a streaming fill (write only) of 4 GiB, with plain and with
        non-temporal stores (the clear's pattern);
      * 2^25 random reads of 16, 32, 64 and 128 bytes from a 6 GiB buffer,
        each writing a 4-byte result (the query's pattern at 32 bytes);
      * 2^24 read-modify-writes of 32 bytes in the same buffer, at random
        lines and in ascending order with random gaps (the segment pass's
        pattern: 16.5 M distinct buckets of 201 M, in bucket order)."""
    import numpy as np
    import torch
    import qsketch_amd as q
    L = q.lib()
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = st.cuda_stream

    def timed(fn):
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            fn()
            e1.record(st)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(ts))

    out = {"copy": {}, "fill": {}, "gather": {}, "rmw": {}}
    for gib in copy_gib:
        src = torch.ones(gib << 28, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        t = timed(lambda: q.check(L.qs_calib_copy(src.data_ptr(), dst.data_ptr(), gib << 30, sp)))
        out["copy"][f"{gib}GiB"] = {"seconds": round(t, 6), "gb_s": round(2 * (gib << 30) / t / 1e9, 1)}
        del src, dst
    dst = torch.empty(1 << 30, dtype=torch.int32, device=dev)  # 4 GiB
    for nt in (0, 1):
        t = timed(lambda: q.check(L.qs_calib_fill(dst.data_ptr(), 4 << 30, 0, nt, sp)))
        out["fill"]["nontemporal" if nt else "plain"] = {"seconds": round(t, 6), "gb_s": round((4 << 30) / t / 1e9, 1)}
    del dst
    if gathers:
        buf = torch.ones(6 << 28, dtype=torch.int32, device=dev)  # 6 GiB
        n = 1 << 25
        res = torch.empty(n, dtype=torch.int32, device=dev)
        for g in gathers:
            t = timed(lambda: q.check(L.qs_calib_gather(buf.data_ptr(), 6 << 30, g, n, 7, res.data_ptr(), sp)))
            out["gather"][f"{g}B"] = {"seconds": round(t, 6), "greads_s": round(n / t / 1e9, 2),
                                      "gb_s_read": round(n * g / t / 1e9, 1),
                                      "gb_s_algorithmic": round(n * (g + 4) / t / 1e9, 1)}
        nr = 1 << 24
        for sorted_ in (0, 1):
            t = timed(lambda: q.check(L.qs_calib_rmw(buf.data_ptr(), 6 << 30, 32, nr, 11, sorted_, sp)))
            out["rmw"]["32B_sorted" if sorted_ else "32B_random"] = {
                "seconds": round(t, 6), "g_rmw_s": round(nr / t / 1e9, 2),
                "gb_s_algorithmic": round(nr * 64 / t / 1e9, 1)}
        del buf, res
    torch.cuda.empty_cache()
    return out


def bench_sf(a):
    """The SF sketch (device_sketch_v42.h) on the measured stream: 2^log2n u32
    keys Zipf(alpha) over 2^log2u, m = 8 counters per bucket split 3 ways (the
    reference driver's choice, src/main.cu:484-485), n by the sizing rule at
    WLF 0.5.  One step = clear, insert (with the update of the largest
    splits), query every distinct key.  Beside it, the oracle on one host
    thread over the first 2^22 keys into a table sized for them."""
    import numpy as np
    import torch
    import qsketch_amd as q
    from oracle import pyoracle as po
    from qsketch_amd import workloads as wl
    torch.cuda.set_device(0)
    N = 1 << a.log2n
    keys = wl.plain_keys(a.alpha, a.log2u, a.log2n)
    uniq = torch.unique(keys)
    counts = torch.empty_like(uniq)
    n = q.number_of_buckets(N, 8, a.wlf)
    sk = q.Sketch_GPU_SF_Host_Sub_Warp(n, 8, 3, 1, 1021, rng_seed=1)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    for _ in range(a.warmup):
        sk.clear()
        sk.insert(keys)
        sk.search(uniq, uniq.numel(), counts)
    ti, tq, tc = [], [], []
    for _ in range(a.steps):
        tc.append(timed(sk.clear))
        ti.append(timed(lambda: sk.insert(keys)))
        tq.append(timed(lambda: sk.search(uniq, uniq.numel(), counts)))
    t_i, t_q, t_c = (float(np.median(x)) for x in (ti, tq, tc))
    # CPU: the oracle, one thread, on a bounded prefix with a table sized for it
    hk = keys[:1 << 22].cpu().numpy().view(np.uint32).copy()
    hu = np.unique(hk)
    o = po.OracleSF(q.number_of_buckets(hk.size, 8, a.wlf), 8, 3, 1021, rng_seed=1)
    t0 = time.perf_counter()
    o.insert(hk)
    t1 = time.perf_counter()
    o.query(hu)
    t2 = time.perf_counter()
    out = {
        "metric": "SF sketch (split counters) insert + query Gkeys/s",
        "value": round((N + uniq.numel()) / (t_i + t_q) / 1e9, 4),
        "unit": "Gkeys/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round((t_c + t_i + t_q) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"SF sketch m=8 times=3, n={n} (WLF {a.wlf}); insert 2^{a.log2n} u32 keys "
                               f"{'zipf' + str(a.alpha) if a.alpha > 0 else 'uniform'} over 2^{a.log2u}, "
                               f"then query the {uniq.numel()} distinct keys",
                   "keys": N, "distinct": int(uniq.numel()), "memory_bytes": sk.memory_bytes()},
        "insert_gkeys_s": round(N / t_i / 1e9, 3), "insert_ms": round(t_i * 1e3, 3),
        "query_gkeys_s": round(uniq.numel() / t_q / 1e9, 3), "query_ms": round(t_q * 1e3, 3),
        "cpu_baseline": {"value": round((hk.size + hu.size) / (t2 - t0) / 1e9, 5), "unit": "Gkeys/s", "cores": 1,
                         "kind": "port",
                         "sample": f"first 2^22 keys of the same stream into a table sized for them, then their "
                                   f"{hu.size} distinct keys queried; oracle on one thread (insert "
                                   f"{hk.size / (t1 - t0) / 1e6:.1f} Mkeys/s with its update)"},
    }
    print(json.dumps(out), flush=True)


def bench_calib(a):
    """Only the ceilings (see ceilings), as one JSON line."""
    import torch
    torch.cuda.set_device(0)
    out = {"metric": "ceilings of the sketch's access patterns on this GPU", "unit": "GB/s"}
    out.update(ceilings())
    out["value"] = out["copy"]["4GiB"]["gb_s"]
    print(json.dumps(out), flush=True)


def learned_compare(log2n, scales, train_steps, reps, n_high, wlf, single_level=True):
    """The learned heavy-hitter form against the plain sketch at equal total
    memory, on the cfg5 stream: 2^log2n u64 keys, Zipf 1.1 over a 2^24-key
    universe, the upper 16 bits of a key clip(round(log2(rank + 1)) +
    N(0, 1.5), 0, 31), the lower 48 bits a permutation of the rank.  The
    network is trained on what a plain sketch estimated for the distinct keys
    of the first 2^26 items (target log2(1 + estimate)); the threshold admits
    at most 2^16 of those keys to the exact table.  Errors over every distinct
    key of the stream against exact counts (sort + reduce-by-key).  Returns
    (rows per memory scale, training record, config)."""
    import numpy as np
    import torch
    import qsketch_amd as q
    from qsketch_amd import accuracy as acc
    from qsketch_amd import learned as lf
    from qsketch_amd import workloads as wl

    N = 1 << log2n
    W = min(N, 1 << 26)
    TOP, TL = 1 << 16, 19
    keys = wl.learned_keys(log2n=log2n)
    n_low0 = q.number_of_buckets(N, 8, wlf)
    wk, tgt = lf.window_targets(keys[:W], n_low0, n_high)
    L0 = q.LearnedSketch(q.make_params(1021, n_high, key_bytes=8, rng_seed=1), table_log2=TL, init_seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = lf.train_router(L0, wk, tgt, steps=train_steps)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    params = L0.get_params()
    thr = lf.pick_threshold(L0, wk, TOP)
    window_routed = int((L0.score(wk) >= thr).sum())
    train_pairs = int(wk.numel())
    L0.destroy()
    del wk, tgt
    uk, uc = torch.unique(keys, return_counts=True)

    def time_it(setup, fn):
        ts = []
        for _ in range(reps):
            setup()
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    rows = []
    for scale in scales:
        n_low = q.largest_prime_at_most(max(1021, int(n_low0 * scale)))
        plain = q.Sketch_GPU_Sub_Warp_Level(n_low, 8, n_high, 32, key_bytes=8, rng_seed=1)
        mem_plain = plain.memory_bytes()
        n_low_l = lf.equal_memory_n_low(n_low, n_high, TL)
        L = q.LearnedSketch(q.make_params(n_low_l, n_high, key_bytes=8, rng_seed=1), table_log2=TL)
        L.set_params(params)
        L.threshold = thr
        mem_learned = L.memory()["total"]
        plain.insert(keys)  # warm-up (scratch allocation)
        L.insert(keys)
        t_plain = time_it(plain.clear, lambda: plain.insert(keys))
        t_learned = time_it(L.clear, lambda: L.insert(keys))
        L.clear()
        n_heavy = L.insert(keys)
        plain.clear()
        plain.insert(keys)
        est_p = plain.search(uk)
        est_l = L.search(uk)
        tq_plain = time_it(lambda: None, lambda: plain.search(uk, uk.numel(), est_p))
        tq_learned = time_it(lambda: None, lambda: L.search(uk, uk.numel(), est_l))
        sbuf = torch.empty(N, dtype=torch.float32, device=keys.device)
        t_score = time_it(lambda: None, lambda: L.score(keys, out=sbuf))
        del sbuf
        mp = acc.metrics(est_p, uc)
        ml = acc.metrics(est_l, uc)
        row = {
            "mem_scale": scale, "n_low_plain": n_low, "n_low_learned": int(n_low_l),
            "memory_plain": mem_plain, "memory_learned": mem_learned,
            "are_plain": mp["are"], "are_learned": ml["are"],
            "are_reduction": round(1 - ml["are"] / mp["are"], 4) if mp["are"] > 0 else None,
            "aae_plain": mp["aae"], "aae_learned": ml["aae"],
            "aae_reduction": round(1 - ml["aae"] / mp["aae"], 4) if mp["aae"] > 0 else None,
            "under_estimates": [mp["under_estimates"], ml["under_estimates"]],
            "insert_gkeys_s_plain": round(N / t_plain / 1e9, 3),
            "insert_gkeys_s_learned": round(N / t_learned / 1e9, 3),
            "insert_throughput_loss": round(1 - t_plain / t_learned, 4),
            "query_gkeys_s_plain": round(uk.numel() / tq_plain / 1e9, 3),
            "query_gkeys_s_learned": round(uk.numel() / tq_learned / 1e9, 3),
            "query_throughput_loss": round(1 - tq_plain / tq_learned, 4),
            "heavy_fraction_of_stream": round(n_heavy / N, 4), "table_keys": int(L.table()[0].size),
            "table_full": L.table_full,
            "mlp_forward_gkeys_s": round(N / t_score / 1e9, 3),
            "mlp_forward_tflops": round(N * 2 * (64 * 64 * 2 + 64) / t_score / 1e12, 2),
        }
        if single_level:
            # the reference's single-level sub-warp sketch (m = 8) at the same bytes
            sl = q.Sketch_GPU_Sub_Warp(q.largest_prime_at_most(mem_plain // 32), 8, key_bytes=8, rng_seed=1)
            sl.insert(keys)
            t_sl = time_it(sl.clear, lambda: sl.insert(keys))
            sl.clear()
            sl.insert(keys)
            msl = acc.metrics(sl.search(uk), uc)
            row["single_level_equal_memory"] = {"n": sl.n, "memory": sl.memory_bytes(), "are": msl["are"],
                                                "aae": msl["aae"], "insert_gkeys_s": round(N / t_sl / 1e9, 3)}
            sl.destroy()
        rows.append(row)
        plain.destroy()
        L.destroy()
    training = {"steps": train_steps, "batch": 4096, "seconds": round(t_train, 3),
                "loss_curve_every_100": [round(x, 5) for x in losses], "threshold": thr,
                "window_keys_routed": window_routed, "train_pairs": train_pairs}
    config = {"workload": f"2^{log2n} u64 keys Zipf 1.1 over 2^24, upper 16 bits = clip(round(log2(rank+1)) "
                          f"+ N(0,1.5), 0, 31); MLP 64-64-64-1 trained on plain-sketch estimates of the first "
                          f"2^26 items; <= 2^16 window keys over the threshold; table 2^{TL} slots",
              "keys": N, "distinct": int(uk.numel())}
    del keys, uk, uc
    return rows, training, config


def learned_accuracy(a) -> dict:
    """The "ARE at equal memory" half of the headline metric, for the default
    line: the cfg5 comparison at full memory."""
    rows, training, config = learned_compare(28, [1.0], a.train_steps, 3, 1 << 18, 0.5, single_level=False)
    r = rows[0]
    return {
        "workload": config["workload"], "keys": config["keys"], "distinct": config["distinct"],
        "are_plain": r["are_plain"], "are_learned": r["are_learned"], "are_reduction": r["are_reduction"],
        "aae_plain": r["aae_plain"], "aae_learned": r["aae_learned"],
        "memory_plain": r["memory_plain"], "memory_learned": r["memory_learned"],
        "insert_gkeys_s_plain": r["insert_gkeys_s_plain"], "insert_gkeys_s_learned": r["insert_gkeys_s_learned"],
        "insert_throughput_loss": r["insert_throughput_loss"],
        "query_gkeys_s_plain": r["query_gkeys_s_plain"], "query_gkeys_s_learned": r["query_gkeys_s_learned"],
        "query_throughput_loss": r["query_throughput_loss"],
        "heavy_fraction_of_stream": r["heavy_fraction_of_stream"],
        "training_seconds": training["seconds"], "train_steps": training["steps"],
        "final_loss": training["loss_curve_every_100"][-1],
    }


def bench_learned(a):
    """Learned heavy-hitter form against the plain sketch at equal total memory
    (several memory budgets); see learned_compare."""
    rows, training, config = learned_compare(a.log2n, [float(x) for x in a.mem_scales.split(",")],
                                             a.train_steps, a.steps, a.n_high, a.wlf)
    main_row = rows[0]
    out = {
        "metric": "learned heavy-hitter form: ARE reduction at equal memory and insert Gkeys/s vs plain",
        "value": main_row["insert_gkeys_s_learned"],
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": 1,
        "ms_per_step": round((1 << a.log2n) / main_row["insert_gkeys_s_learned"] / 1e6, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 keys; bf16 MLP with fp32 accumulation",
        "data": "synthetic",
        "config": dict(config, train_pairs=training["train_pairs"]),
        "training": training,
        "budgets": rows,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
