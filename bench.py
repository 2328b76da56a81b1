#!/usr/bin/env python3
"""Headline benchmark: node-moves/s of a full partition (merge phases, nodal
MCMC sweeps, golden-section search over B) of the N=500k directed DCSBM
(120 planted blocks, 70% of edges inside the block, out-degree power law
-2.5 on [10, 200], about 11.8M edges), started from one block per node,
K=10 merge proposals per block, nodal batches of 4096 nodes, on one MI355X.

One step = one full partition of that graph with the graph resident in HBM;
the timed steps run without kernel timers, and one further partition with
them supplies the roofline and per-kernel figures.  Prints one JSON line (see
the keys below).  --gpus N with N > 1 is refused: this build runs on one GPU.
--config c4 / c5 run the batched and the served workloads instead.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    import usap_amd
    cfg = usap_amd.CONFIGS[name]
    t = time.perf_counter()
    g = usap_amd.config_graph(name)
    log(f"[bench] graph {name}: N={g.N} E={g.E} generated in {time.perf_counter() - t:.1f}s")
    params = dict(init=1, K=10, seed=cfg["seed"], nodal_batch=4096 if g.N >= 100000 else 0,
                  verbose=int(os.environ.get("USAP_VERBOSE", "0")))
    desc = (f"partition of a directed DCSBM, N={g.N}, E={g.E}, {cfg['K']} planted blocks, "
            f"p_in={cfg['p_in']}, out-degree d^-{cfg['gamma']} on [{cfg['dmin']},{cfg['dmax']}], "
            f"identity init (B0=N), K=10, nodal batch "
            f"{params['nodal_batch'] or 'N/4'}")
    return g, params, desc


def cpu_baseline(g, params, phase_log, final_labels, final_B, seconds):
    """The oracle (clean-room CPU restatement at the reference's complexity,
    16 worker threads like the reference's default executor) timed on bounded
    samples of nodal proposals at two states of the same run: after the first
    merge (largest B) and the final partition (smallest B).  The per-phase CPU
    rate is interpolated in log B between the two, and the run's nodal
    proposals are divided by the resulting CPU time."""
    from oracle.oracle import Oracle
    import usap_amd
    threads = min(16, os.cpu_count() or 1)
    eng = usap_amd.Engine(g, **dict(params, verbose=0))
    eng.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    eng.merge_propose(0)
    eng.merge_apply(g.N // 2)
    first_labels, first_B = eng.labels(), eng.B
    eng.close()
    rates = {}
    for tag, lab, B in (("first", first_labels, first_B), ("final", final_labels, final_B)):
        o = Oracle(g, threads=threads, seed=params["seed"])
        o.set_labels(lab, B)
        n, el = 256, 0.0
        while True:
            t = time.perf_counter()
            o.nodal_eval(0, min(n, g.N), 0)
            el = time.perf_counter() - t
            if el > seconds / 4 or n >= g.N:
                break
            n = min(g.N, int(n * max(2.0, (seconds / 4) / max(el, 1e-3))))
        rates[tag] = (B, min(n, g.N) / el, min(n, g.N), el)
        o.close()
        log(f"[bench] cpu sample {tag}: B={B} {min(n, g.N)} proposals in {el:.2f}s "
            f"-> {min(n, g.N) / el:.0f}/s ({threads} threads)")
    (b1, r1, _, _), (b2, r2, _, _) = rates["first"], rates["final"]

    def rate(B):
        if b1 == b2:
            return r1
        f = (math.log(B) - math.log(b2)) / (math.log(b1) - math.log(b2))
        f = min(1.0, max(0.0, f))
        return math.exp(math.log(r2) + f * (math.log(r1) - math.log(r2)))

    props = sum(p["proposals"] for p in phase_log)
    cpu_s = sum(p["proposals"] / rate(p["B"]) for p in phase_log)
    return {
        "value": props / cpu_s,
        "unit": "node-moves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle nodal proposals timed at B={b1} ({rates['first'][2]} proposals, "
                   f"{rates['first'][1]:.0f}/s) and B={b2} ({rates['final'][2]} proposals, "
                   f"{rates['final'][1]:.0f}/s); the run's {props} proposals over {len(phase_log)} "
                   f"phases priced by log-B interpolation of those rates (extrapolated "
                   f"nodal time {cpu_s:.0f}s)"),
    }


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size: under `rocprofv3 --pmc FETCH_SIZE`
    (and, in a second run, WRITE_SIZE) each of its dispatches should read
    (write) nbytes; the ratio scales the same counters of the engine's
    kernels to bytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {reps} copies of {nbytes} bytes")


def measured_traffic(config):
    """HBM bytes per k_nodal launch for this workload, as recorded in the
    counter summary under profiles/ (None when there is none)."""
    path = os.path.join(ROOT, "profiles", f"k_nodal_{config}_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def batched_bench(n_graphs=256, workers=0, replay=0):
    """The batched workload (n_graphs seeded DCSBMs, N log-uniform in
    [500, 50000]) partitioned one graph at a time (the plain form: one
    engine, its launches back to back) and then with several graphs in
    flight (workers from the autotuner unless given).  Both runs must give
    every graph the same labels and MDL."""
    import usap_amd
    t = time.perf_counter()
    graphs, specs = usap_amd.batch_graphs(n=n_graphs, seed=4)
    log(f"[bench] batched workload: {len(graphs)} graphs, {sum(g.N for g in graphs)} nodes, "
        f"{sum(g.E for g in graphs)} edges, generated in {time.perf_counter() - t:.1f}s")
    # reference defaults (SCC init, ref dS); replay: sweeps from captured HIP graphs
    params = [dict(seed=s["seed"], graphs=replay) for s in specs]
    trials = []
    if workers == 0:
        t = time.perf_counter()
        workers, trials = usap_amd.autotune(graphs, params)
        log(f"[bench] autotune: {trials} -> {workers} workers ({time.perf_counter() - t:.1f}s)")
    lab1, st1, b1 = usap_amd.partition_batch(graphs, params, workers=1)
    log(f"[bench] plain: {b1['ms_total']:.0f} ms, {b1['graphs_per_s']:.2f} graphs/s")
    labw, stw, bw = usap_amd.partition_batch(graphs, params, workers=workers)
    log(f"[bench] batched ({workers} in flight): {bw['ms_total']:.0f} ms, "
        f"{bw['graphs_per_s']:.2f} graphs/s")
    same = (b1["failed"] == 0 and bw["failed"] == 0 and
            all(np.array_equal(a, b) for a, b in zip(lab1, labw)) and
            all(x["S"] == y["S"] for x, y in zip(st1, stw)))
    return {
        "metric": "graphs/s",
        "value": round(bw["graphs_per_s"], 3),
        "plain_graphs_per_s": round(b1["graphs_per_s"], 3),
        "speedup_vs_plain": round(bw["graphs_per_s"] / max(b1["graphs_per_s"], 1e-12), 2),
        "workers": workers,
        "autotune": trials,
        "same_mdl_and_labels": bool(same),
        "graphs": len(graphs),
        "nodes": int(sum(g.N for g in graphs)),
        "edges": int(sum(g.E for g in graphs)),
        "ms_plain": round(b1["ms_total"], 1),
        "ms_batched": round(bw["ms_total"], 1),
        "nodal_proposals": int(bw["nodal_proposals"]),
        "mean_nmi": round(float(np.mean([usap_amd.nmi(g.truth, l)
                                         for g, l in zip(graphs, labw)])), 4),
        "workload": (f"{len(graphs)} directed DCSBMs, N log-uniform in [500, 50000], "
                     "round(N^0.4) planted blocks, p_in uniform in [0.55, 0.85], out-degree "
                     "d^-2.5 on [10,100], SCC init, K=10, nodal batch N/4"),
    }


def served_bench(n_graphs=24, rates=(0.5, 1.0, 2.0, 4.0, 8.0), workers=8, interval_s=0.05,
                 check=4, replay=0, final_partition=True):
    """The served stream: n_graphs DCSBMs (N log-uniform in [1000, 200000])
    arriving by a Poisson process at each offered rate, each delivered as 10
    increments of 10% of its shuffled edges, interval_s apart; stage 1
    partitions, later stages resume from the previous partition on the
    device-resident engine.  Per rate: achieved graphs/s and stage latency
    percentiles.  The first `check` graphs are also served alone, and their
    per-stage MDL and final labels must equal those in the mix."""
    from usap_amd import serve as sv
    t = time.perf_counter()
    graphs = sv.served_graphs(n=n_graphs, seed=5)
    log(f"[bench] served workload: {len(graphs)} graphs, {sum(g.N for g in graphs)} nodes, "
        f"{sum(g.E for g in graphs)} edges, generated in {time.perf_counter() - t:.1f}s")
    params = [dict(seed=g.seed, graphs=replay) for g in graphs]
    points, mixes = [], []
    for rate in rates:
        arr = sv.poisson_arrivals(len(graphs), rate, seed=5)
        labs, per, st = sv.serve(graphs, arr, params, stages=10, interval_s=interval_s,
                                 workers=workers, final_partition=final_partition)
        log(f"[bench] served at {rate} graphs/s: {st['graphs_per_s']:.2f} graphs/s, "
            f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, resident <= "
            f"{st['max_resident']}, wall {st['wall_s']:.1f}s")
        points.append({"offered": rate, "graphs_per_s": round(st["graphs_per_s"], 3),
                       "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1),
                       "p50_first_ms": round(st["p50_first_ms"], 1),
                       "p99_first_ms": round(st["p99_first_ms"], 1),
                       "p50_resume_ms": round(st["p50_resume_ms"], 1),
                       "p99_resume_ms": round(st["p99_resume_ms"], 1),
                       "max_resident": st["max_resident"], "failed": st["failed"]})
        mixes.append((labs, per))
    same = True
    for i in range(min(check, len(graphs))):
        lab1, per1, _ = sv.serve([graphs[i]], [0.0], [params[i]], stages=10, workers=1,
                                 final_partition=final_partition)
        for labs, per in mixes:
            same &= bool(np.array_equal(labs[i], lab1[0]))
            same &= [r["S"] for r in per[i]] == [r["S"] for r in per1[0]]
    keep = [p for p in points if p["graphs_per_s"] >= 0.9 * min(p["offered"], len(graphs))]
    last = mixes[-1]
    return {
        "metric": "served graphs/s",
        "value": max((p["graphs_per_s"] for p in keep), default=0.0),
        "points": points,
        "same_alone_and_in_mix": bool(same),
        "checked_alone": min(check, len(graphs)),
        "mean_final_nmi": round(float(np.mean([usap_nmi(g.truth, l)
                                               for g, l in zip(graphs, last[0])])), 4),
        "workers": workers,
        "workload": (f"{len(graphs)} directed DCSBMs, N log-uniform in [1000, 200000], "
                     "round(N^0.4) planted blocks, p_in uniform in [0.55, 0.85], Poisson "
                     f"arrivals, 10 increments of 10% of the shuffled edges {interval_s}s apart; "
                     "stage 1 partition (SCC init), stages 2-9 resume the chain, stage 10 "
                     + ("partitions the whole graph" if final_partition else "resumes")),
    }


def usap_nmi(truth, labels):
    import usap_amd
    return usap_amd.nmi(truth, labels)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graphs", type=int, default=256, help="batched workload size (c4)")
    ap.add_argument("--workers", type=int, default=0, help="graphs in flight (c4; 0 = autotune)")
    ap.add_argument("--replay", action="store_true",
                    help="c4/c5: replay nodal sweeps from captured HIP graphs")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration()
        return
    if args.config == "c5":
        out = served_bench(n_graphs=args.graphs if args.graphs != 256 else 24,
                           workers=args.workers or 8, replay=int(args.replay))
        out.update({"unit": "graphs/s", "n_gpus": 1, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None, "dtype": "int128",
                    "data": "synthetic (seeded directed DCSBM generator)"})
        print(json.dumps(out), flush=True)
        return
    if args.config == "c4":
        out = batched_bench(args.graphs, args.workers, replay=int(args.replay))
        out.update({"unit": "graphs/s", "n_gpus": 1, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None, "dtype": "int128",
                    "data": "synthetic (seeded directed DCSBM generator)"})
        print(json.dumps(out), flush=True)
        return

    import usap_amd
    g, params, desc = workload(args.config)
    eng = usap_amd.Engine(g, **params)
    for w in range(args.warmup):
        t = time.perf_counter()
        eng.partition()
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.2f}s")
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        lab, st = eng.partition()
        stats.append(st)
        log(f"[bench] step {k}: {st['ms_total']:.0f} ms, B={st['B']}, S={st['S']:.1f}, "
            f"phases={st['phases']} sweeps={st['sweeps']} batches={st['batches']}")
    el = time.perf_counter() - t0
    props = sum(s["nodal_proposals"] for s in stats)
    value = props / el
    # kernel times and byte tallies from one more partition, outside the timed
    # region: the per-launch events slow the run down
    eng.profile(1)
    t = time.perf_counter()
    _, pst = eng.partition()
    log(f"[bench] profiled partition: {time.perf_counter() - t:.2f}s")
    prof = eng.profile(0)
    plog = eng.partition_log()
    assert pst["S"] == stats[-1]["S"] and pst["batches"] == stats[-1]["batches"]

    kern = max(prof, key=lambda k: prof[k]["ms"])
    if prof["nodal"]["ms"] > 0:
        kern_roof = "nodal"  # the nodal proposal kernel carries the byte tally
    else:
        kern_roof = kern
    kp = prof[kern_roof]
    achieved = (kp["bytes"] / 1e9) / (kp["ms"] / 1e3) if kp["ms"] > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "k_nodal",
        "achieved": round(achieved, 2),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 5),
        "traffic": measured_traffic(args.config),
        "bytes_per_launch": kp["bytes"] / max(kp["launches"], 1),
        "avg_launch_us": 1e3 * kp["ms"] / max(kp["launches"], 1),
        "launches": kp["launches"],
    }
    out = {
        "metric": "node-moves/s",
        "value": round(value, 1),
        "unit": "node-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int128",
        "data": "synthetic (seeded directed DCSBM generator)",
        "config": {"workload": desc, "N": g.N, "E": int(g.E), "parallelism": "1 GPU"},
        "roofline": roofline,
        "partition": {
            "B": stats[-1]["B"], "S": stats[-1]["S"], "phases": stats[-1]["phases"],
            "sweeps": stats[-1]["sweeps"], "batches": stats[-1]["batches"],
            "nodal_proposals": stats[-1]["nodal_proposals"],
            "accepted_moves": stats[-1]["accepted_moves"],
            "merge_proposals": stats[-1]["merge_proposals"],
            "nmi": round(usap_amd.nmi(g.truth, lab), 5),
            "ms": {k: round(stats[-1][k], 1) for k in
                   ("ms_total", "ms_merge", "ms_carry", "ms_nodal", "ms_entropy", "ms_golden")},
            "rebuilds": stats[-1]["rebuilds"],
        },
        "kernels_ms": {k: round(v["ms"], 2) for k, v in prof.items()},
        "cpu_baseline": None,
    }
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, params, plog, lab, stats[-1]["B"],
                                           args.cpu_seconds)
        out["vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
