#!/usr/bin/env python3
"""Batched multiway evolution on one MI355X: canonical states/sec, served/batched vs plain GPU.

Workload (one "step" = one pass of the hot path over one batch): 4,096 independent multiway
evolutions drawn from seeds 0..4095 (hypergraphrewritingengine_amd/workloads.py:batched_run:
a connected 2-3 edge binary LHS over <= 4 variables, 2-4 RHS edges with <= 2 fresh variables,
a connected random initial state of 3-64 edges, 3-6 steps, a 4,096-state cap per run), all
evolved together from device-resident inputs (hgx_batch_prepare once, outside the timed
region; hgx_batch_evolve per step). The timed region is the device evolution plus the
per-generation host decisions (cap contract); the rate with host inputs (parse + upload inside)
is reported beside it as host_inputs.

Reported beside the value:
  plain     the same requests one at a time through the same library (a bounded sample)
  roofline  the dominant kernel's algorithmic bytes / event-timed duration vs 8 TB/s HBM
  cpu       the reference CPU engine (oracle/_ref/hgref, built from the reference sources)
            timed on this host on a bounded sample, each run to the same generation as the GPU
            (so the same states), with its counts cross-checked against the GPU's.
  others    config 2 (one large evolution), configs 3a/3b (large-state joins), a sampled
            evolution, quotient exploration of the reference's published WPP workload (depths
            5-7, with the raw unfolding replayed beneath it), and ir_audit:
            the host exact (IR) canonicalizer over every raw state of 256 batched runs -- its
            throughput and the WL audit (distinct WL vs distinct IR hashes per run) -- with the
            reference IRCanonicalizer (oracle/_ref/irref) on 20,000 of them, hashes compared.

    python bench.py [--steps K] [--warmup W] [--runs N] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "canonical states/sec + p50/p99 request latency, served/batched vs plain GPU"
HBM_PEAK_GBS = 8000.0


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def cpu_baseline(reqs, gpu_counts, budget_s, threads):
    """The reference engine on the first runs of the batch until budget_s of engine time."""
    from oracle import pyoracle as po

    kind = "reference" if po.ref_available() else "port"
    canon = raw = 0
    secs = 0.0
    n = 0
    mismatch = 0
    t_start = time.perf_counter()
    for req, gc in zip(reqs, gpu_counts):
        if secs >= budget_s or time.perf_counter() - t_start > 3 * budget_s:
            break
        from hypergraphrewritingengine_amd.rules import Request
        same = Request(req.rules, req.init, gc["steps_done"], state_cap=0, flags=req.flags)
        if kind == "reference":
            rc, _ = po.ref_evolve(same, threads=threads, reps=1)
            secs += rc["median_s"]
            c = {"raw_states": rc["raw_states"], "events": rc["events"],
                 "canonical_states": rc["canonical_states"], "causal_triples": rc["causal_triples"],
                 "causal_pairs": rc["causal_pairs"], "branchial_pairs": rc["branchial_pairs"],
                 "fingerprint": rc["fingerprint"]}
        else:
            t = time.perf_counter()
            c = po.evolve(same).counts
            secs += time.perf_counter() - t
        for k, v in c.items():
            if k in gc and gc[k] != v:
                mismatch += 1
                break
        canon += c["canonical_states"]
        raw += c["raw_states"]
        n += 1
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": canon / secs if secs else None, "unit": "canonical states/s",
            "cores": threads if kind == "reference" else 1, "kind": kind,
            "cpu_model": model, "nproc": os.cpu_count(),
            "cores_note": "worker threads are held to the CPU share of one GPU on the measuring host (16), "
                          "whatever nproc reports",
            "sample": f"first {n} of the {len(reqs)} runs, each evolved to the GPU's generation count "
                      f"({raw} raw / {canon} canonical states, {secs:.2f} s engine time)",
            "raw_states_per_s": raw / secs if secs else None, "count_mismatches": mismatch}


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, as recorded from rocprofv3 FETCH_SIZE
    and WRITE_SIZE passes (scaled by a device copy of known size) in the committed profile
    summaries: lines `traffic_per_launch <kernel> <bytes>`."""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc*.md"))):
        for line in open(f):
            m = re.match(r"^traffic_per_launch\s+(\S+)\s+(\d+)", line.strip().strip("`"))
            if m and m.group(1) == kernel:
                return {"bytes": int(m.group(2)), "source": os.path.relpath(f, ROOT)}
    return None


def copy_calibration(n):
    """One device-to-device copy of n bytes (the runtime's copy kernel reads n and writes n
    bytes): the known-size launch that scales FETCH_SIZE / WRITE_SIZE counter readings."""
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    ok = (hip.hipMemset(a, 1, C.c_size_t(n)) == 0 and hip.hipDeviceSynchronize() == 0 and
          hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0 and hip.hipDeviceSynchronize() == 0)
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("calibration copy failed")
    print(f"[bench] copied {n} bytes device-to-device", file=sys.stderr, flush=True)


def dominant(stats):
    name, st = max(stats.items(), key=lambda kv: kv[1]["total_ms"])
    gbs = st["alg_bytes"] / (st["total_ms"] * 1e-3) / 1e9
    return {"kernel": name, "achieved_GBps": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 5),
            "avg_launch_ms": round(st["total_ms"] / st["launches"], 4),
            "alg_bytes_per_launch": st["alg_bytes"] // st["launches"]}


def config2_line(ctx, steps, cpu_steps, threads, no_cpu):
    """One large evolution: rule {{x,y},{x,z}} -> {{x,z},{x,w},{y,w},{z,w}} from {{1,1},{1,1}},
    full multiway with causal and branchial relations."""
    from hypergraphrewritingengine_amd import workloads as W
    ctx.evolve(W.config2(steps=min(steps, 4))).close()
    req = W.config2(steps=steps)
    t = time.perf_counter()
    r = ctx.evolve(req)
    wall = time.perf_counter() - t
    c = r.counts
    r.close()
    ctx.profile(True)
    ctx.evolve(req).close()
    stats = ctx.profile_read()
    ctx.profile(False)
    out = {"steps": steps, "wall_ms": round(wall * 1e3, 3), "raw_states": c["raw_states"],
           "canonical_states": c["canonical_states"], "branchial_pairs": c["branchial_pairs"],
           "causal_pairs": c["causal_pairs"],
           "raw_states_per_s": round(c["raw_states"] / wall, 1),
           "canonical_states_per_s": round(c["canonical_states"] / wall, 1),
           "roofline": dominant(stats)}
    # atomic throughput of the state map: one 64-bit compare-and-swap insert (plus probes) per
    # child state (one per event), all in k_dedup
    dd = stats.get("k_dedup")
    if dd and dd["total_ms"] > 0:
        out["state_map"] = {"inserts": c["events"], "k_dedup_ms": round(dd["total_ms"], 4),
                            "inserts_per_s": round(c["events"] / (dd["total_ms"] * 1e-3), 1)}
    if not no_cpu:
        from oracle import pyoracle as po
        if po.ref_available():
            small = W.config2(steps=cpu_steps)
            rc, _ = po.ref_evolve(small, threads=threads, reps=1)
            g = ctx.evolve(small)
            gc = g.counts
            g.close()
            out["cpu_reference"] = {"steps": cpu_steps, "threads": threads, "s": round(rc["median_s"], 4),
                                    "raw_states_per_s": round(rc["raw_states"] / rc["median_s"], 1),
                                    "counts_match": all(gc[k] == rc[k] for k in
                                                        ("raw_states", "canonical_states", "causal_pairs",
                                                         "branchial_pairs"))}
    return out


def sampled_line(ctx, steps=8, rate=0.25, seed=1):
    """The large evolution's rule thinned by a transition draw (TransitionRate, hgx_sampling):
    eight steps, where the full evolution would hold about 10^10 states. The engine keys each
    draw on the transition's exact hash and ranks (host IR of every frontier state) and expands
    only the survivors."""
    from hypergraphrewritingengine_amd import workloads as W
    req = W.config2(steps=steps)
    req.sampling = {"rate": rate, "seed": seed}
    ctx.evolve(W.config2(steps=4)).close()
    t = time.perf_counter()
    r = ctx.evolve(req)
    wall = time.perf_counter() - t
    c = r.counts
    r.close()
    return {"steps": steps, "transition_rate": rate, "seed": seed, "wall_ms": round(wall * 1e3, 3),
            "raw_states": c["raw_states"], "canonical_states": c["canonical_states"],
            "steps_done": c["steps_done"], "raw_states_per_s": round(c["raw_states"] / wall, 1)}


# The reference's published quotient-exploration timings (paper/tables/t7_gpu.tex:8-10): the WPP
# rule from {{0,1},{0,2}}, exact states, quotient exploration; raw states and events of the
# explored evolution, and its GPU evolve wall (RTX 4090) in ms.
T7_PUBLISHED = {5: (401, 400, 6.2), 6: (3867, 3866, 14.1), 7: (45317, 45316, 378.4)}


def quotient_line(ctx, depths=(5, 6, 7, 8)):
    """Quotient exploration (ExploreFromCanonicalStatesOnly, exact states) of the WPP rule, the
    reference's published GPU workload, and the raw unfolding beneath it replayed on the device
    (hgx_result_unfold): explored states/events against the published ones (depths 5-7; depth 8
    has no published figure), and the walls."""
    from hypergraphrewritingengine_amd import abi
    from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
    rules = [rule_from_lists([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]], 0)]
    flags = 0x1 | 0x2 | 0x8 | 0x100 | abi.HGX_OPT_QUOTIENT
    ctx.evolve(Request(rules, [[0, 1], [0, 2]], 4, flags=flags)).close()
    out = {}
    for d in depths:
        t = time.perf_counter()
        r = ctx.evolve(Request(rules, [[0, 1], [0, 2]], d, flags=flags))
        t_explore = time.perf_counter() - t
        ctx.profile(True)
        t = time.perf_counter()
        u = r.unfold()
        t_unfold = time.perf_counter() - t
        stats = ctx.profile_read()
        ctx.profile(False)
        c = r.counts
        r.close()
        row = {"explored_states": c["raw_states"], "explored_events": c["events"],
               "explore_ms": round(t_explore * 1e3, 3), "unfold_ms": round(t_unfold * 1e3, 3),
               "raw_events": u["raw_events"], "raw_causal_pairs_tr": u["causal_pairs_tr"],
               "raw_branchial_pairs": u["branchial_pairs"], "classes": u["classes"]}
        k = stats.get("unfold_apply")
        if k:
            row["unfold_kernel_ms"] = round(k["total_ms"], 4)
            row["unfold_kernel_alg_GBps"] = round(k["alg_bytes"] / max(k["total_ms"], 1e-9) / 1e6, 2)
        if d in T7_PUBLISHED:
            s0, e0, ms = T7_PUBLISHED[d]
            row["published"] = {"raw_states": s0, "events": e0, "gpu_ms": ms, "gpu": "RTX 4090",
                                 "counts_equal": (c["raw_states"], c["events"]) == (s0, e0),
                                 "explore_speedup": round(ms / (t_explore * 1e3), 2)}
        out[f"steps{d}"] = row
    return out


def config3_line(ctx, name, rule, edges, no_cpu):
    """All matches of one rule in one large random state (device CSR index + seeded join)."""
    ctx.match([rule], 0, edges)
    t = time.perf_counter()
    n, _, dev_ms = ctx.match([rule], 0, edges)
    wall = time.perf_counter() - t
    ctx.profile(True)
    ctx.match([rule], 0, edges)
    stats = ctx.profile_read()
    ctx.profile(False)
    out = {"edges": len(edges), "matches": n, "device_ms": round(dev_ms, 4),
           "wall_ms": round(wall * 1e3, 3), "matches_per_s": round(n / (dev_ms * 1e-3), 1),
           "kernels": {k: {"ms": round(v["total_ms"], 4),
                           "alg_GBps": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 2)}
                       for k, v in stats.items()},
           "roofline": dominant(stats)}
    if not no_cpu:
        from oracle import pyoracle as po
        t = time.perf_counter()
        cn = po.match([rule], 0, edges)
        out["cpu_port"] = {"s": round(time.perf_counter() - t, 4), "threads": 1, "count_match": cn == n}
    return out


def ir_audit_line(ctx, runs, no_cpu):
    """The host exact (IR) canonicalizer over every raw state of a sample of the batched
    workload: its throughput, and the WL audit -- per run, the number of distinct WL hashes
    against the number of distinct IR hashes (equal: WL merged no non-isomorphic states)."""
    import hypergraphrewritingengine_amd as hgx
    from hypergraphrewritingengine_amd import workloads as W
    reqs = [W.batched_run(i, flags=0x3 | 0x8) for i in range(runs)]
    res = ctx.evolve_batch(reqs)
    states, wl_runs = [], []
    for r in res:
        so, eo, vv = r.state_edges()
        st = [[vv[eo[e]:eo[e + 1]].tolist() for e in range(int(so[s]), int(so[s + 1]))] for s in range(len(so) - 1)]
        states.append(st)
        wl_runs.append(r.states()[0])
    flat = [st for run in states for st in run]
    hgx.ir_hash(flat[:64])
    t = time.perf_counter()
    ir = hgx.ir_hash(flat)
    dt = time.perf_counter() - t
    wl_classes = ir_classes = 0
    k = 0
    for run, wl in zip(states, wl_runs):
        wl_classes += len(set(int(h) for h in wl))
        ir_classes += len(set(int(h) for h in ir[k:k + len(run)]))
        k += len(run)
    out = {"runs": runs, "states": len(flat), "host_threads": min(os.cpu_count() or 1, 16),
           "ir_s": round(dt, 4), "ir_states_per_s": round(len(flat) / dt, 1),
           "wl_classes": wl_classes, "ir_classes": ir_classes, "wl_merged_nothing": wl_classes == ir_classes}
    irref = os.path.join(ROOT, "oracle", "_ref", "irref")
    if not no_cpu and os.path.exists(irref):
        # the reference canonicalizer (one thread; it also builds each canonical form) on a sample
        sample = flat[:20000]
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(f"{len(sample)}\n")
            for st in sample:
                f.write(f"{len(st)} " + " ".join(f"{len(e)} " + " ".join(map(str, e)) for e in st) + "\n")
            path = f.name
        try:
            t = time.perf_counter()
            lines = subprocess.run([irref, path], capture_output=True, text=True, check=True).stdout.splitlines()
            rt = time.perf_counter() - t
        finally:
            os.unlink(path)
        ref = [int(x.split()[0]) for x in lines]
        out["cpu_reference"] = {"states": len(sample), "s": round(rt, 3), "threads": 1,
                                "states_per_s": round(len(sample) / rt, 1),
                                "hash_mismatches": sum(int(a) != b for a, b in zip(ir[:len(sample)], ref))}
    return out


def poisson_arrivals(rate, seconds, seed):
    """Arrival times (s) of a Poisson process of `rate` req/s over `seconds`."""
    import random
    rnd = random.Random(seed)
    t, out = 0.0, []
    while True:
        t += rnd.expovariate(rate)
        if t >= seconds:
            return out
        out.append(t)


def serve_window(ctx, rate, seconds, first, seed):
    """Serving under Poisson arrivals: requests W.served_run(first + k) arrive at the k-th
    arrival time, join the next generation launch (continuous batching) and are drained as
    they finish. Latency = completion - arrival, on the host clock. A request arrives in the
    form the C ABI takes (hgx_run_desc arrays, built before the window); its parsing and
    admission are inside the window."""
    from hypergraphrewritingengine_amd import workloads as W
    from hypergraphrewritingengine_amd import abi
    arr = poisson_arrivals(rate, seconds, seed)
    reqs = []
    for k in range(len(arr)):
        r = W.served_run(first + k)
        abi.make_run_desc(r, [])            # requests arrive as the C ABI's arrays
        reqs.append(r)
        if (k + 1) % 50000 == 0:
            print(f"[bench] serve {rate:g} req/s: {k + 1} of {len(arr)} requests built", file=sys.stderr, flush=True)
    n = len(reqs)
    idx_of, lat, done_at = {}, [0.0] * n, [0.0] * n
    canon = raw = 0
    i = done = 0
    steps = 0
    t0 = time.perf_counter()
    while done < n:
        now = time.perf_counter() - t0
        while i < n and arr[i] <= now:
            idx_of[ctx.submit(reqs[i])] = i
            i += 1
        if i == done:                       # idle: wait for the next arrival
            time.sleep(max(0.0, min(arr[i] - now, 0.001)))
            continue
        ctx.serve_step()
        steps += 1
        fin = ctx.drain()
        now = time.perf_counter() - t0
        for rid, res in fin.items():
            k = idx_of.pop(rid)
            c = res.counts
            canon += c["canonical_states"]
            raw += c["raw_states"]
            lat[k] = now - arr[k]
            done_at[k] = now
            res.close()
            done += 1
    wall = max(done_at) if n else 0.0
    return {"rate_rps": rate, "window_s": seconds, "requests": n, "launch_steps": steps,
            "canonical_states_per_s": round(canon / wall, 1) if wall else None,
            "raw_states_per_s": round(raw / wall, 1) if wall else None,
            "p50_ms": round(pct(lat, 50) * 1e3, 3) if n else None,
            "p99_ms": round(pct(lat, 99) * 1e3, 3) if n else None,
            "wall_s": round(wall, 3)}, reqs, arr


def plain_queue(ctx, reqs, arr, budget_s):
    """The same arrivals served one request at a time (each request evolved alone, in arrival
    order, the next starting when the previous finishes): a single-server queue whose service
    times are measured live. Stops after budget_s of wall time; when it stops early the
    latencies cover the requests served so far (a lower bound on the full window's queueing)."""
    t0 = time.perf_counter()
    lat, canon = [], 0
    for r, a in zip(reqs, arr):
        now = time.perf_counter() - t0
        if now > budget_s:
            break
        if now < a:
            time.sleep(a - now)
        out = ctx.evolve(r)
        lat.append(time.perf_counter() - t0 - a)
        canon += out.counts["canonical_states"]
        out.close()
    wall = time.perf_counter() - t0
    return {"canonical_states_per_s": round(canon / wall, 1), "p50_ms": round(pct(lat, 50) * 1e3, 3),
            "p99_ms": round(pct(lat, 99) * 1e3, 3), "wall_s": round(wall, 3),
            "served": len(lat), "of": len(reqs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--runs", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-rates", default="1000,5000",
                    help="Poisson arrival rates (req/s) of the serving windows; '' to skip")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-plain-seconds", type=float, default=8.0)
    ap.add_argument("--no-serve-plain", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="skip the config 2 / 3 lines")
    ap.add_argument("--others", default="config2,config3a,config3b,sampled,quotient,ir_audit",
                    help="which of the other workloads to measure (comma list)")
    ap.add_argument("--config2-steps", type=int, default=6)
    ap.add_argument("--forward", type=int, default=-1, choices=(-1, 0, 1),
                    help="incremental match forwarding: 1 on, 0 off, -1 automatic")
    ap.add_argument("--expand-lanes", type=int, default=0, choices=(0, 16, 32, 64),
                    help="lanes per child in the expansion (HGX_OPTION_EXPAND_LANES; 0 = tuned/automatic)")
    ap.add_argument("--join-lanes", type=int, default=0, choices=(0, 16, 32, 64),
                    help="lanes per parent state in the join (HGX_OPTION_JOIN_LANES; 0 = tuned/automatic)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="skip the graph-replayed passes")
    ap.add_argument("--no-relations", dest="relations", action="store_false",
                    help="skip the relations-materialised pass")
    ap.add_argument("--readback-sample", type=int, default=256)
    ap.add_argument("--autotune", default="auto", choices=("auto", "on", "off"),
                    help="tune launch shapes on the batch before the timed passes (untimed): auto = when "
                         "no tuned shapes were loaded for this device")
    ap.add_argument("--counter-flush", type=int, default=0, metavar="MIB",
                    help="hardware-counter runs: stream MIB MiB through the L2 before and after every "
                         "timed kernel (HGX_OPTION_COUNTER_FLUSH), so each kernel's write-backs are its own")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (a known-size kernel for scaling "
                         "hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the batched path shards whole "
                                   "requests across GPUs with no exchange",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import hypergraphrewritingengine_amd as hgx
    from hypergraphrewritingengine_amd import workloads as W

    ctx = hgx.Context(0)
    ctx.set_forwarding(args.forward)
    ctx.set_expand_lanes(args.expand_lanes)
    ctx.set_join_lanes(args.join_lanes)
    if args.counter_flush:
        ctx.set_counter_flush(args.counter_flush)
    print(f"[bench] {ctx.info}", file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration << 20)
        # a second calibration of known traffic: a fill kernel writing N bytes, then a read-only
        # kernel reading them (k_fill, k_read_calibration)
        ctx.calibration_read(args.copy_calibration << 20)
    reqs = [W.batched_run(i) for i in range(args.runs)]
    batch = ctx.prepare(reqs)               # host-side request records, built once
    tune_s = None
    if args.autotune == "on" or (args.autotune == "auto" and "autotuned" not in ctx.tuning()["source"]
                                 and "tuned (" not in ctx.tuning()["source"]):
        t = time.perf_counter()
        ctx.autotune(batch, reps=2)
        tune_s = round(time.perf_counter() - t, 2)
    tuning = ctx.tuning()
    print(f"[bench] launch shapes: {tuning} (tuning took {tune_s} s)", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        ctx.evolve_batch(batch)
    print("[bench] warmup done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = ctx.evolve_batch(batch)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    counts = [r.counts for r in res]
    canon = sum(c["canonical_states"] for c in counts)
    raw = sum(c["raw_states"] for c in counts)
    value = canon / (ms_per_step / 1000.0)
    print(f"[bench] batched: {ms_per_step:.2f} ms/step, {canon} canonical / {raw} raw states; "
          f"last pass {ctx.last_timing}", file=sys.stderr, flush=True)

    # the same pass recorded into one HIP graph: one graph launch per pass (graph on; the line
    # above is graph off)
    graph = None
    if args.graph:
        t = time.perf_counter()
        rp = ctx.capture(batch)
        cap_s = time.perf_counter() - t
        rp.launch()
        t0g = time.perf_counter()
        for _ in range(args.steps):
            gres = rp.launch()
        g_ms = (time.perf_counter() - t0g) * 1000.0 / args.steps
        gcounts = gres.counts()
        same = all(int(gcounts["canonical_states"][i]) == counts[i]["canonical_states"] and
                   int(gcounts["raw_states"][i]) == counts[i]["raw_states"] for i in range(len(counts)))
        graph = {"ms_per_step": round(g_ms, 3),
                 "canonical_states_per_s": round(canon / (g_ms / 1000.0), 1),
                 "capture_s": round(cap_s, 3), "counts_equal_eager": same,
                 "note": "hgx_batch_capture: the pass recorded once into one HIP graph, replayed with one "
                         "graph launch per step; every kernel reruns, the per-run totals are checked on the "
                         "device against the recording"}
        print(f"[bench] graph: {graph}", file=sys.stderr, flush=True)
        del gres
        rp.close()

    # the same batch handed over as host buffers (parse + upload inside the call)
    t = time.perf_counter()
    ctx.evolve_batch(reqs)
    host_ms = (time.perf_counter() - t) * 1e3
    pcie = {"ms_per_step": round(host_ms, 3), "canonical_states_per_s": round(canon / (host_ms * 1e-3), 1),
            "note": "requests passed as host buffers: parsing and the host-to-device copy of rules "
                    "and initial states included"}

    # roofline of the dominant kernel: one more pass with every launch event-timed
    ctx.profile(True)
    ctx.evolve_batch(batch)
    stats = ctx.profile_read()
    ctx.profile(False)
    dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom["alg_bytes"] / (dom["total_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "avg_launch_ms": round(dom["total_ms"] / dom["launches"], 4),
                "alg_bytes_per_launch": dom["alg_bytes"] // dom["launches"],
                "kernels": {k: {"launches": v["launches"], "ms": round(v["total_ms"], 3),
                                "alg_GBps": round(v["alg_bytes"] / max(v["total_ms"], 1e-9) / 1e6, 2)}
                            for k, v in stats.items()}}

    tr = pmc_traffic(dom_name)
    if tr:
        roofline["traffic"] = tr["bytes"]
        roofline["traffic_source"] = tr["source"]

    # the same batch with the causal and branchial pair lists materialised on the device, and every
    # array of a sample of its results read back (the reference engine always materialises them)
    relations = None
    if args.relations:
        rreqs = [W.batched_run(i, flags=0x7) for i in range(args.runs)]
        rbatch = ctx.prepare(rreqs)
        ctx.evolve_batch(rbatch)
        t = time.perf_counter()
        rres = ctx.evolve_batch(rbatch)
        rel_ms = (time.perf_counter() - t) * 1e3
        rcanon = sum(int(x) for x in rres.counts()["canonical_states"])
        k = min(args.readback_sample, len(rres))
        t = time.perf_counter()
        nread = 0
        for i in range(k):
            r = rres[i]
            h, _, _ = r.states()
            ev = r.events()
            cp, _ = r.causal()
            br = r.branchial()
            nread += len(h) + len(ev) + len(cp) + len(br)
        rb_s = time.perf_counter() - t
        relations = {"ms_per_step": round(rel_ms, 3), "canonical_states_per_s": round(rcanon / (rel_ms * 1e-3), 1),
                     "readback": {"runs": k, "records": nread, "s": round(rb_s, 3),
                                  "records_per_s": round(nread / rb_s, 1) if rb_s else None},
                     "note": "flags causal|branchial|relations: pair lists built on the device; the readback "
                             "sample reads states, events, causal and branchial pairs of the first runs"}
        print(f"[bench] relations: {relations}", file=sys.stderr, flush=True)
        del rres, rbatch

    # plain path: the same requests one at a time, serially
    plain = None
    if not args.no_plain:
        sample = reqs[: args.plain_sample]
        lat = []
        pc = 0
        for r in sample:
            t = time.perf_counter()
            out = ctx.evolve(r)
            lat.append(time.perf_counter() - t)
            pc += out.counts["canonical_states"]
        tot = sum(lat)
        plain = {"canonical_states_per_s": pc / tot, "sample_runs": len(sample),
                 "p50_ms": round(pct(lat, 50) * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3)}
        print(f"[bench] plain: {plain}", file=sys.stderr, flush=True)

    # serving: continuous batching under Poisson arrivals, against the plain one-at-a-time queue
    serving = []
    rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
    for k, rate in enumerate(rates):
        first = 1_000_000 * (k + 1)
        sv, sreqs, sarr = serve_window(ctx, rate, args.serve_seconds, first, seed=k + 1)
        if not args.no_serve_plain:
            sv["plain"] = plain_queue(ctx, sreqs, sarr, args.serve_plain_seconds)
            sv["speedup_vs_plain"] = (round(sv["canonical_states_per_s"] / sv["plain"]["canonical_states_per_s"], 2)
                                      if sv["canonical_states_per_s"] else None)
        serving.append(sv)
        print(f"[bench] serve: {sv}", file=sys.stderr, flush=True)

    # the other measured workloads: one large evolution, and the large-state join
    others = {}
    want = set() if args.no_others else {x.strip() for x in args.others.split(",") if x.strip()}
    threads = min(os.cpu_count() or 1, 16)
    if "config2" in want:
        others["config2"] = config2_line(ctx, args.config2_steps, 4, threads, args.no_cpu)
        print(f"[bench] config2: {others['config2']}", file=sys.stderr, flush=True)
    for name, mk in (("config3a", W.config3a), ("config3b", W.config3b)):
        if name in want:
            rule, edges = mk()
            others[name] = config3_line(ctx, name, rule, edges, args.no_cpu)
            print(f"[bench] {name}: {others[name]}", file=sys.stderr, flush=True)
    if "sampled" in want:
        others["sampled"] = sampled_line(ctx)
        print(f"[bench] sampled: {others['sampled']}", file=sys.stderr, flush=True)
    if "quotient" in want:
        others["quotient"] = quotient_line(ctx)
        print(f"[bench] quotient: {others['quotient']}", file=sys.stderr, flush=True)
    if "ir_audit" in want:
        others["ir_audit"] = ir_audit_line(ctx, 256, args.no_cpu)
        print(f"[bench] ir_audit: {others['ir_audit']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(reqs, counts, args.cpu_seconds, threads)
        # the same engine on one thread, on a shorter sample
        one = cpu_baseline(reqs, counts, args.cpu_seconds / 3, 1)
        cpu["single_thread"] = {k: one[k] for k in ("value", "raw_states_per_s", "sample", "count_mismatches")}
        print(f"[bench] cpu: {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "canonical states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "batched config: 4096 seeded multiway runs, 2-3 edge binary LHS, "
                               "3-64 edge connected inits, 3-6 steps, 4096-state cap per run",
                   "runs": args.runs, "parallelism": "single GPU, one batch per launch sequence",
                   "forwarding": {-1: "automatic", 0: "off", 1: "on"}[args.forward]},
        "raw_states_per_s": round(raw / (ms_per_step / 1000.0), 1),
        "latency_ms": {"batched_p50": round(ms_per_step, 3), "batched_p99": round(ms_per_step, 3),
                       "plain_p50": plain["p50_ms"] if plain else None,
                       "plain_p99": plain["p99_ms"] if plain else None},
        "host_inputs": pcie,
        "plain_gpu": plain,
        "serving": serving,
        "other_workloads": others,
        "speedup_vs_plain": round(value / plain["canonical_states_per_s"], 2) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tuning": {"shapes": tuning, "autotune_s": tune_s},
        "graph": graph,
        "relations": relations,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
