"""Benchmark: exact result pairs/s of the epsilon-grid self-join on one MI355X.

One step = one full device self-join of the workload with its input already
resident in HBM: REORDER + grid build (sort, cells, neighbor runs, tiles) +
counting pass + scan + fill of the CSR neighbor table into a preallocated
buffer.  value = pairs (self-pairs included) / step time.

Prints one JSON line (rank 0) with the roofline of the dominant kernel
(algorithmic bytes / its average HIP-event duration) and a CPU baseline (the
oracle on a bounded query sample, extrapolated to the whole workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 / fp16 MFMA
VALU_PEAK_TFLOPS = {4: 157.3, 8: 78.6}  # MI355X FP32 / FP64 (by DTYPE size)
# VALU issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane-instructions/s
# (non-packed FP32 and FP64 alike; packed FP32 is not used)
VALU_ISSUE_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def valu_ops_per_pair(d, ilp, shortcircuit):
    """VALU lane-instructions of one candidate pair's distance and compare,
    in the accumulation order the parameters select: SEQ = sub + FMA per
    dim; ILP with short-circuit = sub + FMA for each group's first dim,
    sub + mul + add for the others; ILP without = sub + FMA per dim plus
    the fold of the ILP partials."""
    if ilp == 0:
        ops = 2 * d
    elif shortcircuit:
        ops = 3 * d - (d + ilp - 1) // ilp
    else:
        ops = 2 * d + min(ilp, d) - 1
    return ops + 1


def cpu_baseline(x, eps, params, sample_frac, seed=123):
    """Times the CPU oracle on a seeded query sample; extrapolates pairs/s."""
    from oracle import oracle as O
    n = x.shape[0]
    nq = max(1000, int(round(sample_frac * n)))
    nq = min(nq, n)
    rng = np.random.default_rng(seed)
    qids = np.sort(rng.choice(n, nq, replace=False)).astype(np.uint32)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    r = O.join_subset(x, eps, params["k"], qids, ilp=params["ilp"], shortcircuit=params["shortcircuit"],
                      reorder=params["reorder"], nthreads=threads)
    wall = time.perf_counter() - t0
    st = r["stats"]
    t_grid = st[5] * 1e-6
    t_query = st[6] * 1e-6
    pairs_sample = int(r["counts"].sum(dtype=np.uint64))
    scale = n / nq
    t_full = t_grid + t_query * scale
    value = pairs_sample * scale / t_full
    return {"value": value, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": "oracle grid join (full grid build, %d of %d queries = %.2f%%, %d pairs) in %.1f s; "
                      "extrapolated linearly in queries" % (nq, n, 100.0 * nq / n, pairs_sample, wall)}


def _budgeted_queries(n, per_query_s, budget_s, lo=200):
    return int(min(n, max(lo, budget_s / max(per_query_s, 1e-9))))


def cpu_reference_paths(x, eps, budget_s=8.0, seed=321):
    """The reference's CPU paths restated (sequential R-tree, brute force on
    1 core and on all cores), each timed on a seeded query sample sized for
    about budget_s of work and extrapolated linearly in queries.  As in the
    reference, the R-tree's insertion time is reported but not counted."""
    from oracle import oracle as O
    n, d = x.shape
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n).astype(np.uint32)
    threads = min(16, os.cpu_count() or 1)
    out = []

    def entry(name, cores, pairs, nq, secs, extra=""):
        scale = n / nq
        return {"name": name, "value": pairs * scale / (secs * scale), "unit": "pairs/s", "cores": cores,
                "kind": "port", "sample": "%d of %d queries (%.2f%%), %d pairs in %.2f s, extrapolated linearly in "
                "queries%s" % (nq, n, 100.0 * nq / n, pairs, secs, extra)}

    if n * d <= 16_000_000:
        probe = np.sort(perm[:200])
        c, tb, tq = O.cpu_rtree(x, eps, probe)
        nq = _budgeted_queries(n, tq / len(probe), budget_s)
        q = np.sort(perm[:nq])
        c, tb, tq = O.cpu_rtree(x, eps, q)
        out.append(entry("rtree_sequential", 1, int(c.sum(dtype=np.uint64)), nq, tq,
                         "; tree insertion %.2f s not counted" % tb))
    for name, cores in (("bruteforce_1core", 1), ("bruteforce_allcores", threads)):
        probe = np.sort(perm[:64])
        c, t = O.cpu_bruteforce(x, eps, probe, cores)
        nq = _budgeted_queries(n, t / len(probe), budget_s, lo=64)
        q = np.sort(perm[:nq])
        c, t = O.cpu_bruteforce(x, eps, q, cores)
        out.append(entry(name, cores, int(c.sum(dtype=np.uint64)), nq, t))
    return out


EST = {"exact": 0, "learned": 1, "sample": 2, "plain": 0, "nostamp": 0, "linearstamp": 0}

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(workload, kernel_label):
    """HBM bytes per launch of a kernel (read + written), from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy of known size); None when that workload and kernel were not
    measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
        e = tab["workloads"][workload][kernel_label]
        return int(e["read_bytes"] + e["write_bytes"]), e
    except (OSError, KeyError, ValueError):
        return None, None


def copy_probe(nbytes, reps=5):
    """A plain device copy of nbytes (the calibration kernel for the HBM
    byte counters: FETCH_SIZE / WRITE_SIZE per dispatch against nbytes)."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for _ in range(reps):
        ev[0].record()
        dst.copy_(src)
        ev[1].record()
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
    best = min(ms)
    print(json.dumps({"copy_probe_bytes": n * 4, "reps": reps, "best_ms": best,
                      "read_plus_write_GBps": 2 * n * 4 / (best * 1e-3) / 1e9}))


def filter_error(pts, eps, base):
    """One untimed join with the Gram filter's error measurement: the max over
    rechecked pairs of |Gram-form squared distance - FP64 squared distance|,
    relative to eps^2, and the band hit rate (rechecked pairs that are
    neighbors)."""
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        r = sj.join(pts, eps, filter=1, filter_stats=1, estimator=0, **{k: v for k, v in base.items()
                                                                        if k not in ("filter", "estimator")})
        st = r.stats
        return {"max_abs_err_over_eps2": st["filter_max_err"], "rechecked_pairs": st["filter_rechecks"],
                "band_hits": st["filter_hits"],
                "band_hit_rate": st["filter_hits"] / max(1, st["filter_rechecks"]),
                "candidates": st["candidates"],
                "filtered_out_frac": 1.0 - st["filter_rechecks"] / max(1, st["candidates"])}
    finally:
        sj.close()


def run_variant(sj, pts, eps, kw, steps, warmup, total=None):
    """Times `steps` joins (after `warmup`) with output into preallocated buffers."""
    import torch
    n = pts.shape[0]
    dev = pts.device
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)
    if total is None:
        total = sj.join(pts, eps, counts=counts, offsets=offsets, stream=stream.cuda_stream, **kw).total
    ids = torch.empty(max(total, 1), dtype=torch.int32, device=dev)

    def step():
        res = sj.join(pts, eps, out_ids=ids, counts=counts, offsets=offsets, stream=stream.cuda_stream, **kw)
        assert res.total == total
        return res.stats

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        per.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    del ids
    return total, 1e3 * wall / steps, per


def rejoin_rate(pts, eps, kw, total, steps, warmup):
    """Repeated joins of the same points over one grid (gdsj_rejoin_device):
    the counting or slotted pass, scan and fill per step, the grid built
    once.  Reported beside value (which rebuilds the grid every step)."""
    import torch

    from gpu_self_join_amd import SelfJoin
    n = pts.shape[0]
    dev = pts.device
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
    ids = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    sj = SelfJoin(0)
    try:
        sj.join(pts, eps, out_ids=ids, counts=counts, offsets=offsets, stream=stream, **kw)
        for _ in range(warmup):
            sj.rejoin(out_ids=ids, counts=counts, offsets=offsets, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = sj.rejoin(out_ids=ids, counts=counts, offsets=offsets, stream=stream)
            assert r.total == total
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
    finally:
        sj.close()
    return {"ms_per_join": round(ms, 4), "pairs_per_s": total / (ms * 1e-3), "steps": steps,
            "path": "gdsj_rejoin_device: grid, runs and tiles built once, passes repeated"}


def host_path(x, eps, kw, steps=2, warmup=1):
    """The drop-in's host-buffer path (gdsjoin): host points in, the counts
    and the table copied back over PCIe, the library's per-device handle
    reused across calls.  Reported beside value, never as it."""
    from gpu_self_join_amd import _lib, gdsjoin
    p = _lib.default_params(**kw)
    dt = "float" if x.dtype == np.float32 else "double"
    for _ in range(warmup):
        gdsjoin(x, eps, x.shape[1], dt, params=p)
    t0 = time.perf_counter()
    for _ in range(steps):
        counts, table = gdsjoin(x, eps, x.shape[1], dt, params=p)
    wall = (time.perf_counter() - t0) / steps
    return {"value": len(table) / wall, "unit": "pairs/s", "ms_per_join": round(1e3 * wall, 3), "steps": steps,
            "path": "gdsjoin(): host array in, counts and table out over PCIe"}


def bench_batched(args):
    """The batched workload: 256 instances in one job (gdsj_batch_*)."""
    import torch

    from gpu_self_join_amd import BatchJoin, datagen
    dev = torch.device("cuda:0")
    m = datagen.BATCHED_COUNT
    host = [datagen.batched_instance(i)[0] for i in range(m)]
    xs = [torch.from_numpy(x).to(dev) for x in host]
    eps = [datagen.batched_eps(i) for i in range(m)]
    npts = sum(x.shape[0] for x in host)
    coord_bytes = sum(x.size * 4 for x in host)
    results = {}
    total = None
    for vname in (args.variants or "exact,learned").split(","):
        bj = BatchJoin(xs, eps, estimator=EST[vname])
        tune = bj.autotune()
        bj.run()  # capture (learned: exact run + training)
        for _ in range(args.warmup):
            r = bj.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = bj.run()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = bj.stats()
        if total is None:
            total = r.total
        assert r.total == total
        ms = 1e3 * wall / args.steps
        results[vname] = {"ms_per_step": round(ms, 4), "pairs_per_s": total / (ms * 1e-3), "tuning": tune,
                          "graph_ms": round(st["ms_last_run"], 4), "stats": st}
        bj.close()
    best = min(results, key=lambda v: results[v]["ms_per_step"])
    r = results[best]
    t = r["tuning"]
    # dominant kernel: the persistent counting pass (the autotuner's HIP-event
    # time of the chosen configuration); it reads every coordinate and writes N counts
    kb = coord_bytes + 4 * npts
    achieved = kb / (t["ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_batch_join<count pass>", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("batched", "k_batch_join<count pass>")[0],
            "alg_bytes_per_launch": kb, "avg_launch_ms": round(t["ms"], 4)}
    out = {
        "metric": "exact result pairs/s (batched grid self-joins, %s)" % best,
        "value": r["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded batched generator, 256 instances)",
        "config": {"workload": "batched", "instances": m, "n_total": npts, "total_pairs": total,
                   "estimator": best, "tuning": t},
        "variants": {k: {kk: vv for kk, vv in v.items() if kk != "stats"} | {"stats": v["stats"]}
                     for k, v in results.items()},
        "roofline": roof,
    }
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_batched(host, eps)
    print(json.dumps(out))


def cpu_baseline_batched(host, eps, stride=32):
    """The oracle on every stride-th instance (whole joins); pairs/s of the sample."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    pairs = 0
    t0 = time.perf_counter()
    idx = list(range(0, len(host), stride))
    for i in idx:
        x = host[i]
        r = O.join(x, eps[i], k=min(x.shape[1], 6), ilp=8, shortcircuit=1, reorder=1, with_ids=False,
                   nthreads=threads)
        pairs += int(r["counts"].sum(dtype=np.uint64))
    wall = time.perf_counter() - t0
    return {"value": pairs / wall, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": "oracle grid join of instances %s (%d pairs) in %.1f s" % (idx, pairs, wall)}


PHASES = ("ms_grid", "ms_estimate", "ms_count", "ms_scan", "ms_compact", "ms_fill", "ms_sort", "ms_train",
          "ms_total")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gmm6d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-path", action="store_true", help="skip timing the PCIe-inclusive gdsjoin() path")
    ap.add_argument("--cpu-sample", type=float, default=0.01)
    ap.add_argument("--variants", default=None,
                    help="estimators to time (exact = counting pass + fill; learned = MLP-sized single pass)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--param", action="append", default=[],
                    help="override a gdsj_params field that does not change the result, e.g. queryreorder=0")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the join shards by query cells; "
              "multi-GPU runs are not part of it)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.workload == "batched":
        return bench_batched(args)

    import torch

    from gpu_self_join_amd import SelfJoin, datagen

    name = args.workload
    w = datagen.WORKLOADS[name]
    x = datagen.generate(name)
    eps = datagen.workload_eps(name)
    params = datagen.workload_params(name)
    n, d = x.shape
    dev = torch.device("cuda:0")
    pts = torch.from_numpy(x).to(dev)
    base = dict(k=params["k"], ilp=params["ilp"], shortcircuit=params["shortcircuit"], reorder=params["reorder"],
                stamp=params["stamp"], queryreorder=params["queryreorder"])
    for item in args.param:
        key, val = item.split("=")
        if key in ("k", "ilp", "reorder", "shortcircuit"):
            raise SystemExit("--param %s would change the workload's result contract" % key)
        base[key] = int(val)
    esz = np.dtype(w["dtype"]).itemsize
    if args.variants is None:
        args.variants = "exact,nostamp,linearstamp,learned" if base.get("stamp") else "exact,learned"
    results = {}
    kws = {}
    total = None
    for vname in args.variants.split(","):
        kw = dict(base, estimator=EST[vname])
        kws[vname] = kw
        sj = SelfJoin(0)
        steps, warmup = args.steps, args.warmup
        if vname == "learned":
            kw["train"] = 0  # trains on the first join's counts, then sizes slots
        if vname == "linearstamp":
            # unicomp with the lower-linear-id ownership (LINEARSTAMP)
            kw["stamp"] = 2
        if vname == "nostamp":
            # the exact passes over the full neighborhood (STAMP only changes
            # which side evaluates a pair, not the table)
            kw["stamp"] = 0
        if vname == "plain":
            # the plain grid join without the MFMA filter (the reference's
            # method on this hardware): few steps, it is the slow reference point
            kw["filter"] = 0
            steps, warmup = min(steps, 3), min(warmup, 1)
        total, ms, per = run_variant(sj, pts, eps, kw, steps, warmup, total)
        sj.close()
        phase = {k: float(np.mean([s[k] for s in per])) for k in PHASES}
        last = per[-1]
        results[vname] = {"ms_per_step": ms, "pairs_per_s": total / (ms * 1e-3), "phase_ms": phase, "stats": last}
    best = min((v for v in results if v != "plain"), key=lambda v: results[v]["ms_per_step"], default="plain")
    r = results[best]
    phase = r["phase_ms"]
    s = r["stats"]
    # dominant kernel and its algorithmic bytes per launch: every pass reads
    # the points once (N*D*s); the counting pass writes N counts, the slotted
    # pass N counts + 5 bytes per pair of slot, the fill pass 4 bytes per pair
    kb = {"count": n * d * esz + 4 * n, "fill": n * d * esz + 4 * total}
    kms = {"count": phase["ms_count"], "fill": phase["ms_fill"]}
    kname = {"count": "k_join<count pass>", "fill": "k_join<fill pass>"}
    if best in ("learned", "sample"):
        kb["count"] = n * d * esz + 4 * n + 5 * total
        kname["count"] = "k_join<slotted pass>"
        kb["compact"] = 9 * total
        kms["compact"] = phase["ms_compact"]
        kname["compact"] = "k_compact"
    dom = max(kms, key=kms.get)
    achieved = kb[dom] / (kms[dom] * 1e-3) / 1e9
    traffic, tsrc = measured_traffic(name, kname[dom])
    roof = {"bound": "hbm", "kernel": kname[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes_per_launch": kb[dom], "avg_launch_ms": round(kms[dom], 4)}
    if tsrc:
        roof["traffic_detail"] = {"read_bytes": tsrc["read_bytes"], "write_bytes": tsrc["write_bytes"],
                                  "source": "profiles/hbm_traffic.json", "kernel": tsrc["kernel"]}
    if s["filter_rechecks"] == 0 and dom != "compact":
        # the plain pass evaluates every candidate pair: a subtract and an FMA
        # per dim (3 D flops), against the FP32 / FP64 vector peak
        vflops = 3.0 * d * s["candidates"]
        vpeak = VALU_PEAK_TFLOPS[esz]
        vt = vflops / (kms[dom] * 1e-3) / 1e12
        ops = valu_ops_per_pair(d, base["ilp"], base["shortcircuit"]) * float(s["candidates"])
        it = ops / (kms[dom] * 1e-3) / 1e12
        roof["valu"] = {"achieved": round(vt, 3), "peak": vpeak, "unit": "TFLOP/s", "frac": round(vt / vpeak, 4),
                        "alg_flops_per_launch": vflops, "candidate_pairs": s["candidates"],
                        "issue": {"achieved": round(it, 3), "peak": round(VALU_ISSUE_PEAK_TOPS, 2),
                                  "unit": "T lane-ops/s", "frac": round(it / VALU_ISSUE_PEAK_TOPS, 4),
                                  "ops_per_pair": valu_ops_per_pair(d, base["ilp"], base["shortcircuit"])}}
    if s["filter_rechecks"] > 0 and dom != "compact":
        # the MFMA Gram filter decides the pass: algorithmic work 2*D flops per
        # candidate pair (C = queries x candidates), against the dense
        # bf16/fp16 MFMA peak; the HBM figure is kept beside it
        flops = 2.0 * d * s["candidates"]
        tf = flops / (kms[dom] * 1e-3) / 1e12
        hbm = dict(roof)
        roof = {"bound": "mfma", "kernel": kname[dom].replace("k_join", "k_gram"), "achieved": round(tf, 3),
                "peak": MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / MFMA_PEAK_TFLOPS, 5),
                "traffic": measured_traffic(name, kname[dom].replace("k_join", "k_gram"))[0],
                "alg_flops_per_launch": flops, "avg_launch_ms": round(kms[dom], 4),
                "hbm": {k: hbm[k] for k in ("achieved", "peak", "unit", "frac", "alg_bytes_per_launch")}}
    variants = {}
    for vname, vr in results.items():
        st = vr["stats"]
        variants[vname] = {"ms_per_step": round(vr["ms_per_step"], 4), "pairs_per_s": vr["pairs_per_s"],
                           "phase_ms": {k: round(v, 4) for k, v in vr["phase_ms"].items()}}
        variants[vname]["filter"] = {"rechecks": st["filter_rechecks"], "hits": st["filter_hits"],
                                     "candidates": st["candidates"]}
        if vname in ("learned", "sample"):
            variants[vname]["estimator"] = {k: st[k] for k in ("est_predicted", "est_capacity", "est_overflow_tiles",
                                                                "est_tiles", "est_mean_rel_err", "est_max_rel_err")}
    out = {
        "metric": "exact result pairs/s (grid self-join, %s)" % best,
        "value": r["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if esz == 4 else "f64",
        "data": "synthetic (seeded %s generator)" % name,
        "config": {"workload": name, "n": n, "dims": d, "eps": eps, "total_pairs": total,
                   "pairs_per_point": total / n, **base, "estimator": best},
        "variants": variants,
        "grid": {"nonempty_cells": s["n_nonempty_cells"], "tiles": s["n_tiles"], "candidates": s["candidates"]},
        "speedup_vs_plain": (results["plain"]["ms_per_step"] / r["ms_per_step"]) if "plain" in results else None,
        # the join passes alone (grid resident): counting or slotted pass, scan,
        # compaction and fill (the grid build is in ms_per_step, not here)
        "t_join_ms": round(sum(phase[k] for k in ("ms_estimate", "ms_count", "ms_scan", "ms_compact", "ms_fill")), 4),
        "pairs_per_s_join": total / (sum(phase[k] for k in ("ms_estimate", "ms_count", "ms_scan", "ms_compact",
                                                            "ms_fill")) * 1e-3),
        "filter_error": filter_error(pts, eps, base) if d >= 16 else None,
        "roofline": roof,
        "host_path": None if args.no_host_path else host_path(x, eps, kws[best]),
        "rejoin": rejoin_rate(pts, eps, kws[best], total, args.steps, args.warmup),
    }
    if args.no_cpu_baseline:
        out["cpu_baseline"] = None
    else:
        refs = cpu_reference_paths(x, eps)
        grid = cpu_baseline(x, eps, params, args.cpu_sample)
        grid["name"] = "grid_oracle_allcores"
        # the headline baseline: the reference's sequential R-tree path where
        # it was timed, else the fastest restated CPU path
        main_ = next((r for r in refs if r["name"] == "rtree_sequential"), None)
        out["cpu_baseline"] = dict(main_ or max(refs + [grid], key=lambda r: r["value"]))
        out["cpu_baselines"] = refs + [grid]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
