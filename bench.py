#!/usr/bin/env python3
"""Benchmark of the HDBSCAN hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload mix30_1m_d12]

A step is one pass of the path over the resident workload: NN-Descent kNN
graph (K=32, SAMPLE_NUM=16) -> hub counts -> anti-hubs -> core distances
(mpts=32) -> mutual-reachability graph -> Boruvka MST -> sorted MST edges.
The points are uploaded to HBM before the timed region.  Prints one JSON line.

Legs (recall@32 measured on 10,000 sampled queries against an exact GPU
brute force; a leg counts only at recall >= --target):
  plain       random init, NND_ITERATION=6, join_keep 4 (the library
              defaults); its kernels supply stage_ms, kernel_ms and roofline
  plain_best  random init at its fastest (join_keep, iterations)
  learned     learned init (8 networks and window 512 at d <= 16, else 6 and
              256; trained once on a 200k subsample, the training time
              reported apart) at its fastest (join_keep, iterations)
value = points / second of the learned leg (of the plain leg when the
learned leg is skipped or misses the target).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mix30_1m_d12")
    ap.add_argument("--iterations", type=int, default=6)
    ap.add_argument("--distance-mode", choices=("fp32", "bf16"), default="fp32",
                    help="bf16: BF16 MFMA candidate tiles with exact FP32 re-rank (the 2M x 128 variant)")
    ap.add_argument("--blocked", type=int, default=0, help="blocked all-pairs kNNG with this many blocks")
    ap.add_argument("--queries", type=int, default=10000)
    ap.add_argument("--cpu-sample", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="time the plain (random-init) form only")
    ap.add_argument("--learned-subsample", type=int, default=200000)
    ap.add_argument("--learned-tables", type=int, default=None,
                    help="bucket tables (networks) of the learned init (default: by dimension)")
    ap.add_argument("--learned-window", type=int, default=None,
                    help="bucket-order window per point and table (default: by dimension)")
    ap.add_argument("--target", type=float, default=0.95, help="recall@32 a timed form must reach")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, three device-to-device copies of this size (counter calibration)")
    return ap.parse_args()


def cpu_baseline(X, sample_n):
    """The oracle's CPU path (FP64 exact kNN, hub counts, anti-hubs, core,
    MRD CSR, Kruskal, stable edge sort) on the first sample_n points."""
    from oracle import oracle as O
    Xs = np.ascontiguousarray(X[:sample_n])
    n = len(Xs)
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ids, d2 = O.bruteforce_fp64(Xs, np.arange(n, dtype=np.int32), 32)
    L = np.empty((n, 32), O.NND_ELEMENT)
    L["label"] = ids
    L["distance"] = d2.astype(np.float32)
    counts = O.hub_counts(L)
    ah = O.antihubs(L, counts)
    core = O.core(Xs, L, 32)
    ni, nl, w = O.mrd_csr(Xs, L, core, ah)
    edges, _ = O.kruskal(ni, nl, w)
    O.sort_edges(edges)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "points/s", "cores": cores, "kind": "port",
            "seconds": round(dt, 3),
            "sample": f"first {n} points of the workload; FP64 exact kNN (k=32, all points as queries) + hub "
                      f"counts + anti-hubs + core + MRD CSR + Kruskal + edge sort, C oracle with OpenMP"}


def recall_at_32(ctx, Xd, lists, q, gt_ids):
    """recall@32 of device lists on query ids q against ground-truth ids."""
    from gpu_hdbscan_amd import api
    rows = api.gather_rows(ctx, lists, q)
    lab = rows["label"]
    hit = sum(len(np.intersect1d(lab[i], gt_ids[i])) for i in range(len(q)))
    return hit / (len(q) * 32.0)


def time_leg(ctx, step, steps, warmup):
    """W untimed steps, then K timed steps on the library stream; returns
    (ms per step, the last result, HIP-event timers, NN-Descent stats)."""
    for _ in range(warmup):
        step()
    ctx.synchronize()
    ctx.timer_enable(True)
    ctx.timer_reset()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    timers = ctx.timer_read()
    stats = ctx.nnd_stats()
    ctx.timer_enable(False)
    return elapsed * 1e3 / steps, res, timers, stats


def best_config(run, check, keeps, hi, target):
    """For each join_keep, the fewest iterations (1..hi) whose lists reach the
    recall target (one run per count); returns [(keep, iterations, recall)]
    of the keeps that reach it."""
    found = []
    for keep in keeps:
        for t in range(1, hi + 1):
            run(keep, t)
            r = check()
            if r >= target:
                found.append((keep, t, r))
                break
    return found


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs the single-GPU path only (--gpus {a.gpus} requested); the multi-GPU sharding is a "
              "design note, not built", file=sys.stderr)
        sys.exit(2)
    from gpu_hdbscan_amd import api, datasets, learned

    X = datasets.make(a.workload)
    n, d = X.shape
    ctx = api.Context(0)
    Xd = ctx.to_device(X)
    bufs = {"lists": ctx.empty((n, 32), api.NND_ELEMENT), "counts": ctx.empty((n,), np.int32),
            "core": ctx.empty((n,), np.float32), "edges": ctx.empty((n - 1,), api.MST_EDGE)}
    mode = 1 if a.blocked else 0
    rng = np.random.default_rng(1234)
    q = rng.choice(n, size=min(a.queries, n), replace=False).astype(np.int32)
    gt_ids, _ = api.bruteforce_knn(ctx, Xd, q, 32)

    def check():
        return recall_at_32(ctx, Xd, bufs["lists"], q, gt_ids)

    # plain form: random init, the reference's NND_ITERATION
    p_plain = api.default_params(nnd_iteration=a.iterations, distance_mode=1 if a.distance_mode == "bf16" else 0)

    def plain_step(prm=p_plain):
        return api.pipeline(ctx, Xd, params=prm, knng_mode=mode, blocks=max(a.blocked, 1), keep_on_device=True,
                            outputs=bufs)

    ms_plain, res_plain, timers, stats = time_leg(ctx, plain_step, a.steps, a.warmup)
    recall_plain = check()

    legs = {"plain": {"iterations": a.iterations, "ms_per_step": round(ms_plain, 3),
                      "points_per_s": n / (ms_plain * 1e-3), "recall_at_32": round(recall_plain, 4)}}
    value, ms_step, recall, init_used = n / (ms_plain * 1e-3), ms_plain, recall_plain, "random"

    if not a.no_learned and not a.blocked:
        # each form at its fastest (join_keep, iterations) that still meets
        # the recall target; every candidate is timed over the same steps
        def fastest(step_for, found):
            best = None
            for keep, t, _ in found:
                prm = api.default_params(nnd_iteration=t, join_keep=keep)
                ms, _, _, _ = time_leg(ctx, lambda: step_for(prm), a.steps, 1)
                r = check()
                if r >= a.target and (best is None or ms < best[0]):
                    best = (ms, keep, t, r)
            return best

        pb = fastest(plain_step, best_config(lambda k, t: plain_step(api.default_params(nnd_iteration=t,
                                                                                          join_keep=k)),
                                             check, (2, 3, 4), 12, a.target))
        if pb:
            legs["plain_best"] = {"join_keep": pb[1], "iterations": pb[2], "ms_per_step": round(pb[0], 3),
                                  "points_per_s": n / (pb[0] * 1e-3), "recall_at_32": round(pb[3], 4)}
        # learned form: networks trained on the GPU from plain NN-Descent over a
        # seeded subsample, then learned initial lists + fewer iterations
        model = learned.LearnedInit(ctx, d, seed=0, tables=a.learned_tables,
                                    window=a.learned_window).fit(X, n_sub=min(a.learned_subsample, n))
        ctx.synchronize()

        def learned_step(prm):
            return learned.pipeline_learned(ctx, Xd, model, params=prm, outputs=bufs)

        lb = fastest(learned_step, best_config(lambda k, t: learned_step(api.default_params(nnd_iteration=t,
                                                                                              join_keep=k)),
                                               check, (2, 3, 4), 8, a.target))
        if lb:
            legs["learned"] = {"join_keep": lb[1], "iterations": lb[2], "ms_per_step": round(lb[0], 3),
                               "points_per_s": n / (lb[0] * 1e-3), "recall_at_32": round(lb[3], 4),
                               "tables": len(model.handles), "window": model.window,
                               "train_s": round(model.train_seconds, 3),
                               "train_subsample": min(a.learned_subsample, n),
                               "final_loss": round(float(model.final_loss), 5),
                               "points_per_s_incl_training": n / (lb[0] * 1e-3 + model.train_seconds)}
            legs["learned_vs_plain"] = round(ms_plain / lb[0], 3)
            if pb:
                legs["learned_vs_plain_best"] = round(pb[0] / lb[0], 3)
            value, ms_step, recall, init_used = n / (lb[0] * 1e-3), lb[0], lb[3], "learned"
        model.close()

    # dominant kernel of the plain leg: the local join, timed by HIP events
    # on the library stream
    join_ms, join_launches = timers.get("nnd_join", (0.0, 0))
    iters = a.iterations
    per_launch_ms = join_ms / max(join_launches, 1)
    # algorithmic bytes of one join launch: per point its sample bookkeeping
    # (24 B), per candidate its vector, its id and its list bound (4d + 8 B),
    # per proposal (target, key) = 12 B; stats are summed over the iterations
    # of the last step
    cand = stats["candidates"] / iters
    props = stats["proposals"] / iters
    join_bytes = n * 24 + cand * (4 * d + 8) + props * 12
    achieved = join_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "join_traffic_mix30_1m_d12.json")
    if a.workload == "mix30_1m_d12" and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    stage_ms = {k: round(v[0] / a.steps, 3) for k, v in sorted(timers.items()) if k.startswith("stage_")}
    kern_ms = {k: round(v[0] / a.steps, 3) for k, v in sorted(timers.items()) if not k.startswith("stage_")}
    out = {
        "metric": "kNNG->MST points/sec at recall@32>=0.95, learned-init vs plain, 1 GPU",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 tiles + f32" if a.distance_mode == "bf16" else "f32",
        "data": "synthetic",
        "config": {"workload": a.workload, "points": n, "dim": d, "k": 32, "sample_num": 16,
                   "nnd_iteration": legs.get("learned", {}).get("iterations", iters) if init_used == "learned"
                   else iters, "mpts": 32, "init": init_used,
                   "knng": f"blocked x{a.blocked}" if a.blocked else "single NN-Descent",
                   "join_keep": legs.get("learned", {}).get("join_keep", int(p_plain.join_keep))
                   if init_used == "learned" else int(p_plain.join_keep)},
        "recall_at_32": round(recall, 4),
        "recall_ok": bool(recall >= a.target),
        "recall_queries": int(len(q)),
        "legs": legs,
        "mst_edges": int(res_plain.n_edges),
        "components": int(res_plain.components),
        "csr_edges": int(res_plain.csr_edges),
        "stage_ms_plain": stage_ms,
        "kernel_ms_plain": kern_ms,
        "roofline": {"kernel": "k_nnd_join", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": int(join_bytes),
                     "avg_launch_ms": round(per_launch_ms, 4)},
        "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, min(a.cpu_sample, n))
    print(json.dumps(out), flush=True)
    if a.copy_calibration_mib > 0:
        nb = a.copy_calibration_mib << 20
        src, dst = ctx.empty((nb,), np.uint8), ctx.empty((nb,), np.uint8)
        for _ in range(3):
            ctx._check(ctx.lib.hdb_memcpy_d2d(ctx.handle, dst.ptr, src.ptr, nb))
        ctx.synchronize()
    ctx.close()


if __name__ == "__main__":
    main()
