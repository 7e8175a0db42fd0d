"""Benchmark: exact result pairs/s of the epsilon-grid self-join on one MI355X.

One step = one full device self-join of a workload with its input already
resident in HBM: REORDER + grid build (sort, cells, neighbor runs, tiles) +
counting (or estimator-sized slotted) pass + scan + fill of the CSR neighbor
table into a preallocated buffer.  value = pairs (self-pairs included) /
step time.

The default run measures the 2,000,000-point 6-D Gaussian mixture (gmm6d)
as the headline line and adds a `workloads` block with the other four
workloads (uniform2d, aniso32d, lowid64d, batched) at reduced step counts,
so one invocation covers every configuration.  `--workload NAME` measures
one workload alone.

Algorithmic work of one join: B = N*D*s + 4*N + 4*P bytes (points read once,
uint32 counts and uint32 neighbor ids written; s = sizeof(DTYPE)) and, for
the MFMA Gram filter, F = 2*D*C flops (C = candidate pairs).  Every roofline
fraction is B (or F) over the dominant kernel's average HIP-event duration
against the MI355X peak; `hbm_frac_join` is B over the whole join (grid
resident).  CPU baselines (the oracle and the restated reference CPU paths)
are timed on bounded query samples and extrapolated; they are reported, not
targets.
"""
import argparse
import json
import os
import subprocess
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

T_START = time.perf_counter()


def log(msg):
    print("[bench %6.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


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


def alg_bytes(n, d, esz, pairs):
    """Algorithmic bytes of one join: points once, counts and ids written."""
    return n * d * esz + 4 * n + 4 * pairs


# ------------------------------------------------------------ CPU baselines --
def cpu_baseline(x, eps, params, sample_frac, seed=123):
    """Times the CPU oracle on a seeded query sample; extrapolates pairs/s."""
    from oracle import oracle as O
    n = x.shape[0]
    nq = max(200, int(round(sample_frac * n)))
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
    return {"name": "grid_oracle_allcores", "value": value, "unit": "pairs/s", "cores": threads, "kind": "port",
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


def cpu_baselines(x, eps, params, budget_s, grid_frac):
    """Headline CPU baseline (the sequential R-tree where it was timed, else
    the fastest restated path) and the full list."""
    refs = cpu_reference_paths(x, eps, budget_s=budget_s)
    grid = cpu_baseline(x, eps, params, grid_frac)
    main_ = next((r for r in refs if r["name"] == "rtree_sequential"), None)
    return dict(main_ or max(refs + [grid], key=lambda r: r["value"])), refs + [grid]


# ------------------------------------------------------------- machine info --
def machine_info(copy_bytes=1 << 30):
    """CPU model and core count of this host, and a measured device-copy rate
    (the best of five copies of copy_bytes: read + write bytes / time)."""
    import torch
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    n = copy_bytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for _ in range(5):
        ev[0].record()
        dst.copy_(src)
        ev[1].record()
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
    best = min(ms)
    del src, dst
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "gpu": torch.cuda.get_device_name(0),
            "copy_probe": {"bytes": n * 4, "best_ms": round(best, 4),
                           "read_plus_write_GBps": round(2 * n * 4 / (best * 1e-3) / 1e9, 1)}}


def copy_probe(nbytes, reps=5):
    """A plain device copy of nbytes (the calibration kernel for the HBM
    byte counters: FETCH_SIZE / WRITE_SIZE per dispatch against nbytes)."""
    info = machine_info(nbytes)
    print(json.dumps(info["copy_probe"]))


# ------------------------------------------------------------ measurements --
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
    # the two C calls of gdsjoin timed apart: the join (points in, counts out)
    # and the table copy into a fresh array (as gdsjoin allocates it)
    import ctypes as C
    L = _lib.load()
    flat = np.ascontiguousarray(x.ravel())
    cnt = np.zeros(x.shape[0], dtype=np.uint32)
    f = L.GDSJoinPy if dt == "float" else L.GDSJoinPyDouble
    ev = C.c_float(np.float32(eps)) if dt == "float" else C.c_double(eps)
    t1 = time.perf_counter()
    f(flat.ctypes.data, x.shape[0], ev, x.shape[1], 0, cnt.ctypes.data)
    t2 = time.perf_counter()
    tab = np.empty(int(cnt.sum(dtype=np.uint64)), dtype=np.uint32)
    L.copyResultIntoPythonArray(tab.ctypes.data, tab.size)
    t3 = time.perf_counter()
    return {"value": len(table) / wall, "unit": "pairs/s", "ms_per_join": round(1e3 * wall, 3), "steps": steps,
            "ms_join_call": round(1e3 * (t2 - t1), 3), "ms_table_copy": round(1e3 * (t3 - t2), 3),
            "table_bytes": int(tab.nbytes),
            "path": "gdsjoin(): host array in, counts and table out over PCIe (a ring of pinned buffers: one DMA stream in chunk order, up to 8 host copy threads)"}


def estimator_holdout(name, x, eps, base):
    """The learned estimator out of sample: trained on the workload's first
    join, then frozen (train = 2) and used to size joins it has not seen
    (another seed of the generator, and epsilon x 0.9 / x 1.1 on the same
    points), each beside the reference's sampling estimator on the same
    join.  Per join: mean / max relative error per tile, overflowed tiles,
    slot capacity over pairs."""
    import torch

    from gpu_self_join_amd import SelfJoin, datagen
    sj = SelfJoin(0)
    ss = SelfJoin(0)
    out = {}
    try:
        pts = torch.from_numpy(x).cuda()
        sj.join(pts, eps, estimator=EST["learned"], train=0, **base)  # exact passes + training
        other = torch.from_numpy(datagen.generate(name, seed=datagen.WORKLOAD_SEEDS[name] + 100)).cuda()
        cases = [("seed+100", other, eps), ("eps_x0.9", pts, float(np.float32(eps * 0.9))),
                 ("eps_x1.1", pts, float(np.float32(eps * 1.1)))]
        for cname, p, e in cases:
            rec = {}
            for est, h, extra in (("learned", sj, {"train": 2}), ("sample", ss, {})):
                r = h.join(p, e, estimator=EST[est], **extra, **base)
                st = r.stats
                rec[est] = {"mean_rel_err": round(st["est_mean_rel_err"], 4),
                            "max_rel_err": round(st["est_max_rel_err"], 3),
                            "overflow_tiles": st["est_overflow_tiles"], "tiles": st["est_tiles"],
                            "overflow_rate": st["est_overflow_tiles"] / max(1, st["est_tiles"]),
                            "capacity_over_pairs": round(st["est_capacity"] / max(1, r.total), 3),
                            "ms_estimate": round(st["ms_estimate"], 4), "ms_total": round(st["ms_total"], 4)}
            out[cname] = rec
        del other
    finally:
        sj.close()
        ss.close()
    return out


PHASES = ("ms_grid", "ms_estimate", "ms_count", "ms_scan", "ms_compact", "ms_fill", "ms_sort", "ms_train",
          "ms_total")
JOIN_PHASES = ("ms_estimate", "ms_count", "ms_scan", "ms_compact", "ms_fill")


def bench_workload(name, args, variants, steps, warmup, cpu_budget, cpu_frac, extras=True):
    """One workload: every variant timed, the best one's roofline, filter
    error (dims >= 16), CPU baselines; returns the record."""
    import torch

    from gpu_self_join_amd import SelfJoin, datagen

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
        if key in ("k", "ilp", "reorder", "shortcircuit", "ilp_fold"):
            raise SystemExit("--param %s would change the workload's result contract" % key)
        base[key] = int(val)
    esz = np.dtype(w["dtype"]).itemsize
    results = {}
    kws = {}
    total = None
    for vname in variants:
        kw = dict(base, estimator=EST[vname])
        kws[vname] = kw
        sj = SelfJoin(0)
        st_, wu = steps, warmup
        if vname == "learned":
            kw["train"] = 0  # trains on the first join's counts, then sizes slots
        if vname == "linearstamp":
            kw["stamp"] = 2  # unicomp with the lower-linear-id ownership (LINEARSTAMP)
        if vname == "nostamp":
            kw["stamp"] = 0  # the full neighborhood (STAMP only changes which side evaluates a pair)
        if vname == "plain":
            # the plain grid join without the MFMA filter on the reference's
            # dims (the reference's method on this hardware): few steps, it is
            # the slow reference point
            kw["filter"] = 0
            kw["index"] = 1
            st_, wu = min(steps, 2 if d < 64 else 1), 0
        total, ms, per = run_variant(sj, pts, eps, kw, st_, wu, total)
        sj.close()
        phase = {k: float(np.mean([s[k] for s in per])) for k in PHASES}
        results[vname] = {"ms_per_step": ms, "pairs_per_s": total / (ms * 1e-3), "phase_ms": phase,
                          "stats": per[-1], "steps": st_}
        log("%s %s: %.3f ms/step" % (name, vname, ms))
    best = min((v for v in results if v != "plain"), key=lambda v: results[v]["ms_per_step"], default="plain")
    r = results[best]
    phase = r["phase_ms"]
    s = r["stats"]
    B = alg_bytes(n, d, esz, total)
    # dominant kernel: the counting (or slotted) pass, the compaction or the
    # fill, by average HIP-event duration
    kms = {"count": phase["ms_count"], "fill": phase["ms_fill"], "compact": phase["ms_compact"]}
    kname = {"count": "k_join<count pass>", "fill": "k_join<fill pass>", "compact": "k_compact"}
    if best in ("learned", "sample"):
        kname["count"] = "k_join<slotted pass>"
    dom = max(kms, key=kms.get)
    gram = s["filter_rechecks"] > 0 and dom != "compact"
    if gram:
        kname[dom] = kname[dom].replace("k_join", "k_gram")
    achieved = B / (kms[dom] * 1e-3) / 1e9
    traffic, tsrc = measured_traffic(name, kname[dom])
    roof = {"bound": "hbm", "kernel": kname[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "alg_bytes_per_launch": B, "avg_launch_ms": round(kms[dom], 4)}
    if tsrc:
        roof["traffic_detail"] = {"read_bytes": tsrc["read_bytes"], "write_bytes": tsrc["write_bytes"],
                                  "source": "profiles/hbm_traffic.json", "kernel": tsrc["kernel"]}
    if not gram and dom != "compact":
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
    if gram:
        # the MFMA Gram filter decides the pass: 2*m flops per candidate pair,
        # m = the operands' dims (D, or the principal axes they are taken on)
        m = int(s.get("gram_axes", 0)) or d
        flops = 2.0 * m * s["candidates"]
        tf = flops / (kms[dom] * 1e-3) / 1e12
        hbm = dict(roof)
        roof = {"bound": "mfma", "kernel": kname[dom], "achieved": round(tf, 3), "peak": MFMA_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tf / MFMA_PEAK_TFLOPS, 5), "traffic": traffic,
                "alg_flops_per_launch": flops, "candidate_pairs": s["candidates"], "operand_dims": m,
                "avg_launch_ms": round(kms[dom], 4),
                "hbm": {k: hbm[k] for k in ("achieved", "peak", "unit", "frac", "alg_bytes_per_launch")}}
    variants_out = {}
    for vname, vr in results.items():
        st = vr["stats"]
        variants_out[vname] = {"ms_per_step": round(vr["ms_per_step"], 4), "pairs_per_s": vr["pairs_per_s"],
                               "steps": vr["steps"],
                               "phase_ms": {k: round(v, 4) for k, v in vr["phase_ms"].items()},
                               "candidates": st["candidates"]}
        if st["filter_rechecks"]:
            variants_out[vname]["filter"] = {"rechecks": st["filter_rechecks"], "hits": st["filter_hits"]}
        if vname in ("learned", "sample"):
            variants_out[vname]["estimator"] = {
                k: st[k] for k in ("est_predicted", "est_capacity", "est_overflow_tiles", "est_tiles",
                                   "est_mean_rel_err", "est_max_rel_err")}
            variants_out[vname]["estimator"]["overflow_rate"] = st["est_overflow_tiles"] / max(1, st["est_tiles"])
            variants_out[vname]["estimator"]["capacity_over_pairs"] = st["est_capacity"] / max(1, total)
    t_join = sum(phase[k] for k in JOIN_PHASES)
    rec = {
        "metric": "exact result pairs/s (grid self-join, %s)" % best,
        "value": r["pairs_per_s"],
        "unit": "pairs/s",
        "ms_per_step": round(r["ms_per_step"], 4),
        "steps": r["steps"],
        "dtype": "f32" if esz == 4 else "f64",
        "config": {"workload": name, "n": n, "dims": d, "eps": eps, "total_pairs": total,
                   "pairs_per_point": total / n, **base, "estimator": best},
        "variants": variants_out,
        "grid": {"nonempty_cells": s["n_nonempty_cells"], "tiles": s["n_tiles"], "candidates": s["candidates"],
                 "index": "principal axes" if s.get("index_used") == 2 else "reference dims"},
        "speedup_vs_plain": (results["plain"]["ms_per_step"] / r["ms_per_step"]) if "plain" in results else None,
        # the join passes alone (grid resident): counting or slotted pass,
        # scan, compaction and fill (the grid build is in ms_per_step)
        "t_join_ms": round(t_join, 4),
        "pairs_per_s_join": total / (t_join * 1e-3),
        "alg_bytes_per_join": B,
        "hbm_frac_join": round(B / (t_join * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "filter_error": filter_error(pts, eps, base) if d >= 16 else None,
        "roofline": roof,
    }
    if extras:
        rec["host_path"] = None if args.no_host_path else host_path(x, eps, kws[best])
        rec["rejoin"] = rejoin_rate(pts, eps, kws[best], total, steps, warmup)
    del pts
    torch.cuda.empty_cache()
    if args.no_cpu_baseline:
        rec["cpu_baseline"] = None
    else:
        log("%s: CPU baselines" % name)
        rec["cpu_baseline"], rec["cpu_baselines"] = cpu_baselines(x, eps, params, cpu_budget, cpu_frac)
    return rec, x, eps, base


def bench_batched(args, steps, warmup, variants):
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
    for vname in variants:
        bj = BatchJoin(xs, eps, estimator=EST[vname])
        tune = bj.autotune()
        bj.run()  # capture (learned: exact run + training)
        for _ in range(warmup):
            r = bj.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = bj.run()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = bj.stats()
        if total is None:
            total = r.total
        assert r.total == total
        ms = 1e3 * wall / steps
        results[vname] = {"ms_per_step": round(ms, 4), "pairs_per_s": total / (ms * 1e-3), "tuning": tune,
                          "graph_ms": round(st["ms_last_run"], 4), "stats": st}
        bj.close()
        log("batched %s: %.3f ms/step" % (vname, ms))
    best = min(results, key=lambda v: results[v]["ms_per_step"])
    r = results[best]
    t = r["tuning"]
    B = coord_bytes + 4 * npts + 4 * total
    # dominant kernel: the persistent counting pass (the autotuner's HIP-event
    # time of the chosen configuration), against the whole batch's B
    achieved = B / (t["ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_batch_join<count pass>", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic("batched", "k_batch_join<count pass>")[0],
            "alg_bytes_per_launch": B, "avg_launch_ms": round(t["ms"], 4)}
    rec = {
        "metric": "exact result pairs/s (batched grid self-joins, %s)" % best,
        "value": r["pairs_per_s"], "unit": "pairs/s", "ms_per_step": r["ms_per_step"], "steps": steps,
        "dtype": "f32",
        "config": {"workload": "batched", "instances": m, "n_total": npts, "total_pairs": total,
                   "estimator": best, "tuning": t},
        "variants": {k: {kk: vv for kk, vv in v.items() if kk != "stats"} | {"stats": v["stats"]}
                     for k, v in results.items()},
        "alg_bytes_per_join": B,
        "roofline": roof,
    }
    del xs
    torch.cuda.empty_cache()
    rec["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_batched(host, eps)
    return rec


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
    return {"name": "grid_oracle_allcores", "value": pairs / wall, "unit": "pairs/s", "cores": threads,
            "kind": "port", "sample": "oracle grid join of instances %s (%d pairs) in %.1f s" % (idx, pairs, wall)}


def headline(rec, args, n_gpus=1):
    """The contract's fields around a workload record."""
    out = {"metric": rec["metric"], "value": rec["value"], "unit": rec["unit"], "n_gpus": n_gpus,
           "steps": rec["steps"], "warmup": args.warmup, "ms_per_step": rec["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": rec["dtype"],
           "data": "synthetic (seeded %s generator)" % rec["config"]["workload"]}
    out.update({k: v for k, v in rec.items() if k not in out})
    return out


SECONDARY = {
    # workload: (variants, steps cap, CPU budget per restated path (s), oracle query fraction)
    "uniform2d": ("exact,nostamp", 10, 3.0, 0.01),
    "aniso32d": ("learned,exact,plain", 5, 3.0, 0.002),
    "lowid64d": ("learned,plain", 3, 3.0, 0.0005),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    help="all (gmm6d headline + the other workloads) or one of uniform2d, gmm6d, aniso32d, "
                         "lowid64d, batched")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-path", action="store_true", help="skip timing the PCIe-inclusive gdsjoin() path")
    ap.add_argument("--cpu-sample", type=float, default=0.01)
    ap.add_argument("--variants", default=None,
                    help="estimators to time (exact = counting pass + fill; learned = MLP-sized single pass; "
                         "sample = the reference's sampling estimator sizing the single pass)")
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
        rec = bench_batched(args, args.steps, args.warmup, (args.variants or "exact,learned").split(","))
        print(json.dumps(headline(rec, args)))
        return
    if args.workload != "all":
        name = args.workload
        variants = args.variants or ("exact,nostamp,linearstamp,learned" if name == "uniform2d" else
                                     "exact,learned,sample")
        rec, _, _, _ = bench_workload(name, args, variants.split(","), args.steps, args.warmup, 8.0,
                                      args.cpu_sample)
        print(json.dumps(headline(rec, args)))
        return

    # default: the gmm6d headline and every other workload in one line
    info = machine_info()
    log("machine %s" % json.dumps(info))
    rec, x, eps, base = bench_workload("gmm6d", args, (args.variants or "exact,learned,sample").split(","),
                                       args.steps, args.warmup, 8.0, args.cpu_sample)
    out = headline(rec, args)
    out["machine"] = info
    log("gmm6d: estimator out of sample")
    out["estimator_holdout"] = estimator_holdout("gmm6d", x, eps, base)
    del x
    work = {}
    for name, (variants, cap, budget, frac) in SECONDARY.items():
        log("workload %s" % name)
        r2, _, _, _ = bench_workload(name, args, variants.split(","), min(args.steps, cap),
                                     min(args.warmup, 2), budget, frac, extras=name == "uniform2d")
        work[name] = r2
    log("workload batched")
    work["batched"] = bench_batched(args, min(args.steps, 5), min(args.warmup, 1), ["exact", "learned"])
    out["workloads"] = work
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
