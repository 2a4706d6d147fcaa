#!/usr/bin/env python3
"""Benchmark of the Cas-OFFinder search path on one MI355X.

A step is one full search of the workload over the genome resident in HBM:
queries submitted -> fused finder+comparer kernels (one launch per pattern
group) -> device sort -> compact records in host memory.  Genome
generation/upload and text formatting are not in the step and are reported
separately.

Default workload (mixed_pam_2000): the synthetic 3.1 Gbp genome (seed 1, 24
contigs, GC 41%, ~5% N, 50% soft-masked), 2,000 guides of four Cas9 PAMs
(seed 2) with thresholds 3-6 and 25% DNA/RNA bulges up to 1, one reference
input file per (PAM, D, R) group, searched batched (the metric's value) and
plain (one finder launch, then one comparer launch per compare key, the
reference's launch model).  The 100-guide NGG and the 20-guide bulge
workloads are reported beside it, then the continuous-batching server on a
Poisson request stream and the CPU baseline.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cas-offinder_amd"))

import casoff  # noqa: E402
import workloads  # noqa: E402

METRIC = "guide×Gbp/s bit-exact, batched vs plain on 1 GPU; serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
CU, CLOCK_HZ = 256, 2.4e9  # VALU issue peak: one wave64 instruction per CU-cycle
# Algorithmic HBM bytes of one batched launch: every launch (one pattern
# group) reads the genome's three bit planes once, 0.375 B per position
# (2-bit base + 1-bit non-ACGT plane); candidates, windows and the key index
# stay on chip (LDS / registers).  roofline.achieved = these bytes over the
# dominant kernel's average launch time, measured here with HIP events around
# each launch on the library's stream (cof_launch_info) over the timed steps.
# The kernels are bound by VALU issue (integer bit work, no dense
# contraction): roofline.valu gives the dominant kernel's executed VALU
# instructions per launch (rocprofv3 SQ_INSTS_VALU, committed summary of the
# same workload) over the issue peak of the same launch time.
BYTES_PER_POSITION = 0.375


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    "genome_3g_ngg": "3.1 Gbp synthetic genome (seed 1, 24 contigs), pattern 20N+NGG, 100 guides, mm<=4",
    "bulge_3g_nrg": "3.1 Gbp synthetic genome (seed 1), 20N+NRG, DNA bulge 2 / RNA bulge 1, 20 guides, mm<=4",
    "mixed_pam_2000": "3.1 Gbp synthetic genome (seed 1), 2000 mixed-PAM guides (seed 2; SpCas9 NGG 40%, SaCas9 "
                      "NNGRRT 30%, StCas9 NNAGAAW 20%, NmCas9 NNNNGMTT 10%; mm 3-6; 25% with DNA/RNA bulge <= 1), "
                      "one reference input per (PAM, D, R) group: 16 groups",
}


def workload_texts(name):
    if name == "genome_3g_ngg":
        return [workloads.genome_3g_ngg("synthetic")]
    if name == "bulge_3g_nrg":
        return [workloads.bulge_3g_nrg("synthetic")]
    if name == "mixed_pam_2000":
        return [t for _, t in workloads.mixed_pam_2000("synthetic")]
    raise SystemExit("unknown workload " + name)


def n_guides_of(texts):
    return sum(len([l for l in t.split("\n")[2:] if l.strip()]) for t in texts)


def profile_kernel(workload, kernel):
    """The committed rocprofv3 figures of one kernel instance on a workload
    (profiles/*.md, HTML comments `<!-- kernel {json} -->`): average
    duration, FETCH_SIZE / WRITE_SIZE bytes (scaled by the copy calibration)
    and SQ_INSTS_VALU per launch."""
    import glob
    import re
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for m in re.finditer(r"<!-- kernel (\{.*\}) -->", open(p).read()):
            d = json.loads(m.group(1))
            if d.get("workload") == workload and d.get("kernel") == kernel:
                return d, os.path.relpath(p, ROOT)
    return None, None


def calibrate_copy(nbytes):
    """A device copy of known size (torch elementwise kernel) for scaling the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = nbytes // 4
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.add(x, 1, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 3}), flush=True)


def time_steps(ctx, qs, mode, steps, warmup):
    """Warmup searches (the first batched one runs the autotuner's trials and
    grows the hit buffer), then `steps` timed searches.  Returns seconds per
    step, kernel ms per step, launches per step, the last stats and result,
    and per kernel instance the summed launch ms and launch count of the
    timed steps."""
    tune_ms = 0.0
    for _ in range(warmup):
        ctx.search(qs, mode)
        tune_ms += ctx.stats()["t_tune_ms"]
    time_steps.tune_ms = tune_ms
    kernel_ms, launches, per_kernel, collect = 0.0, 0, {}, 0.0
    parts = {"sort_ms": 0.0, "copy_ms": 0.0, "decode_ms": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.search(qs, mode)
        st = ctx.stats()
        kernel_ms += st["t_fused_ms"] if mode == casoff.MODE_BATCHED else st["t_finder_ms"] + st["t_comparer_ms"]
        launches += st["n_launches"]
        collect += st["t_collect_ms"]
        for k in parts:
            parts[k] += st["t_" + k]
        if mode == casoff.MODE_BATCHED:
            for rec in ctx.launch_info():
                k = per_kernel.setdefault(rec["kernel"], {"ms": 0.0, "launches": 0})
                k["ms"] += rec["ms"]
                k["launches"] += 1
    t1 = time.perf_counter()
    time_steps.collect_ms = collect / steps
    time_steps.collect_parts = {k: round(v / steps, 3) for k, v in parts.items()}
    return (t1 - t0) / steps, kernel_ms / steps, launches / steps, st, r, per_kernel


def roofline(workload, per_kernel, steps, genome_len):
    """The dominant kernel instance (largest summed launch time over the timed
    steps) against HBM, with its VALU issue rate beside it."""
    kernel, rec = max(per_kernel.items(), key=lambda kv: kv[1]["ms"])
    total_ms = sum(v["ms"] for v in per_kernel.values())
    avg_ms = rec["ms"] / rec["launches"]
    nbytes = int(BYTES_PER_POSITION * genome_len)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    prof, src = profile_kernel(workload, kernel)
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": int(prof["fetch_bytes"] + prof["write_bytes"]) if prof else None,
           "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": round(avg_ms, 4),
           "launches_per_step": rec["launches"] / steps,
           "share_of_kernel_time": round(rec["ms"] / total_ms, 3), "profile": src,
           "note": "algorithmic bytes = 0.375 B x genome positions (three bit planes read once per launch); "
                   "avg_launch_ms from HIP events around each launch over the timed steps"}
    if prof:
        issue = prof["valu_insts"] / (CU * avg_ms * 1e-3 * CLOCK_HZ)
        out["valu"] = {"insts_per_launch": prof["valu_insts"], "issue_per_cu_cycle": round(issue, 3),
                       "peak_issue_per_cu_cycle": 1.0, "profile_avg_launch_ms": prof["avg_us"] / 1e3}
    return out


def run_workload(ctx, name, steps, warmup, plain_steps, with_plain=True):
    texts = workload_texts(name)
    qs = [ctx.query(t) for t in texts]
    n_guides = n_guides_of(texts)
    total = ctx.info()[0]
    log("workload: %s (%d queries, %d guides)" % (name, len(qs), n_guides))
    s_b, k_b, l_b, st_b, r_b, per_kernel = time_steps(ctx, qs, casoff.MODE_BATCHED, steps, warmup)
    hits_b = r_b.records()
    launches = [{k: rec[k] for k in ("kernel", "sig", "variant", "candidates")} | {"ms": round(rec["ms"], 3)}
                for rec in ctx.launch_info()]
    log("batched: %.2f ms/step (kernels %.2f ms), %d hits" % (s_b * 1e3, k_b, len(hits_b)))
    res = {"value": round(n_guides * total / 1e9 / s_b, 2), "ms_per_step": round(s_b * 1e3, 3), "guides": n_guides,
           "queries": len(qs), "kernel_ms_per_step": round(k_b, 3), "collect_ms_per_step": round(time_steps.collect_ms, 3),
           "collect_parts": time_steps.collect_parts,
           "launches_per_step": l_b, "hits": len(hits_b),
           "candidate_sites_per_s": st_b["n_pairs"] / s_b,
           "counters": {"candidates": st_b["n_candidates"], "pairs": st_b["n_pairs"], "hits": len(hits_b),
                        "base_comparisons": st_b["n_base_cmp"]},
           "autotune": {"trial_ms_in_warmup": round(time_steps.tune_ms, 2),
                        "timed_steps_launch": {0: "eager launches", 1: "captured HIP graphs",
                                               2: "replayed HIP graphs"}[int(st_b["graph"])]},
           "launches": launches,
           "roofline": roofline(name, per_kernel, steps, total)}
    if with_plain:
        s_p, k_p, l_p, st_p, r_p, _ = time_steps(ctx, qs, casoff.MODE_PLAIN, plain_steps, 0)
        hits_p = r_p.records()
        log("plain: %.2f ms/step (kernels %.2f ms), %d hits" % (s_p * 1e3, k_p, len(hits_p)))
        same = bool(len(hits_b) == len(hits_p) and (hits_b == hits_p).all())
        res["plain"] = {"value": round(n_guides * total / 1e9 / s_p, 2), "ms_per_step": round(s_p * 1e3, 3),
                        "steps": plain_steps, "launches_per_step": l_p, "bit_exact_vs_batched": same}
        res["batched_over_plain"] = round(s_p / s_b, 2)
        if st_p["t_finder_ms"] > 0:
            # plain path's finder (finder_count_kernel + tile-count scan +
            # finder_write_kernel): 0.375 B read per scanned position + a
            # (locus u32, flag u8) entry written per locus with a PAM
            fbytes = 0.375 * st_p["n_finder_positions"] + 5 * st_p["n_loci"]
            fach = fbytes / (st_p["t_finder_ms"] * 1e-3) / 1e9
            res["plain"]["finder_roofline"] = {
                "kernel": "finder_count_kernel + finder_write_kernel", "bound": "hbm",
                "algorithmic_bytes_per_step": int(fbytes), "ms_per_step": round(st_p["t_finder_ms"], 4),
                "achieved": round(fach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(fach / HBM_PEAK_GBS, 4)}
    # cold step: fresh query objects (parse, tables and indexes built and
    # uploaded, launches captured), tuner decisions of the warmup kept
    t0 = time.perf_counter()
    fresh = [ctx.query(t) for t in texts]
    r = ctx.search(fresh, casoff.MODE_BATCHED)
    res["cold_step_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    assert r.count() == len(hits_b)
    return texts, res


def host_cores():
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the host sets it (the GPU host gives one GPU's
    process a share of its cores)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(ctx, texts, target_s=15.0):
    """The oracle (CPU restatement of the reference's finder/comparer, one
    position / one candidate per iteration, OpenMP) over every query of the
    workload on a prefix of the first contig, scaled to ~target_s of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = host_cores()
    name, pos, clen, _ = ctx.contigs()[0]
    sample = min(clen, 1_000_000)
    with tempfile.TemporaryDirectory() as td:
        for _ in range(3):
            gp = os.path.join(td, "sample.fa")
            with open(gp, "w") as f:
                f.write(">" + name + "\n")
                seq = ctx.text(pos, sample)
                for i in range(0, len(seq), 60):
                    f.write(seq[i:i + 60] + "\n")
            secs, cands, glen = 0.0, 0, 0
            for i, t in enumerate(texts):
                ip = os.path.join(td, "in%d.txt" % i)
                with open(ip, "w") as f:
                    f.write(gp + "\n" + t.split("\n", 1)[1])
                res = oracle.time_search(ip, threads)
                secs += res["seconds"]
                cands += res["candidates"]
                glen = res["genome_len"]
            if secs >= target_s / 3 or sample >= clen:
                break
            sample = min(clen, int(sample * min(4.0, target_s / max(secs, 1e-3))))
    n_guides = n_guides_of(texts)
    return {"value": n_guides * (glen / 1e9) / secs, "unit": "guide×Gbp/s", "cores": threads, "kind": "port",
            "sample": "oracle (reference-algorithm CPU restatement, OpenMP) over all %d queries (%d guides) on the "
                      "first %d bp of %s: %.2f s, %d candidates" % (len(texts), n_guides, sample, name, secs, cands)}


def serve_bench(ctx, n_requests, qps, seed=3, verify=True):
    """Poisson request stream (workloads.serving_stream) through the
    continuous-batching server; latency is submit -> records in host memory."""
    reqs = workloads.serving_stream(n_requests, qps, "synthetic", seed=seed)
    queries = [[ctx.query(t) for _, t in groups] for _, groups, _ in reqs]
    n_guides = sum(n for _, _, n in reqs)
    srv = casoff.Server(ctx, mem_budget_bytes=8 << 30, max_batch_queries=64)
    tickets = []
    t0 = time.perf_counter()
    for (arrival, _, _), qs in zip(reqs, queries):
        wait = arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        tickets.append(srv.submit(qs))
    done = [srv.poll(t, wait=True) for t in tickets]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.stop()
    lat = sorted(ms for _, ms in done)
    pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
    out = {"requests": n_requests, "qps_offered": qps, "guides": n_guides,
           "p50_ms": round(pct(0.50), 2), "p99_ms": round(pct(0.99), 2), "max_ms": round(lat[-1], 2),
           "wall_s": round(wall, 3), "guide_x_gbp_per_s": round(n_guides * ctx.info()[0] / 1e9 / wall, 2),
           "batches": st["batches"], "lane_hits": st["lane_hits"], "lane_grows": st["lane_grows"],
           "admission_waits": st["admission_waits"]}
    if st["batches"]:
        b = st["batches"]
        out["per_batch"] = {"requests": round(st["requests"] / b, 2), "launches": round(st["launches"] / b, 2),
                            "search_ms": round(st["t_search_ms"] / b, 3), "plan_ms": round(st["t_plan_ms"] / b, 3),
                            "kernel_ms": round(st["t_kernel_ms"] / b, 3), "split_ms": round(st["t_split_ms"] / b, 3)}
    if verify:
        same_alone, same_plain, checked_plain = 0, 0, 0
        for (res, _), qs, (_, _, n) in zip(done, queries, reqs):
            got = res.records()
            alone = ctx.search(qs, casoff.MODE_BATCHED).records()
            same_alone += int(len(got) == len(alone) and bool((got == alone).all()))
            if n <= 8:
                plain = ctx.search(qs, casoff.MODE_PLAIN).records()
                checked_plain += 1
                same_plain += int(len(got) == len(plain) and bool((got == plain).all()))
        out["bit_exact_vs_alone"] = "%d/%d" % (same_alone, n_requests)
        out["bit_exact_vs_plain"] = "%d/%d (requests of <= 8 guides)" % (same_plain, checked_plain)
    return out


def serve_capacity(ctx, n_requests, levels, p99_ms, plain_requests):
    """Served capacity: the serving stream at each offered rate in `levels`
    (at least n_requests requests, about 2 s of arrivals); capacity is the highest achieved guide×Gbp/s whose p99 latency is within
    p99_ms.  Beside it, the plain path serving the first plain_requests
    requests one after another (no queueing: its rate is an upper bound of
    what plain serving sustains)."""
    sweep = []
    for qps in levels:  # about 2 s of arrivals at each rate, so that the rate is sustained
        r = serve_bench(ctx, max(n_requests, int(2 * qps)), qps, verify=False)
        log("serving %.0f QPS: p50 %.1f ms, p99 %.1f ms, %.0f guide×Gbp/s" %
            (qps, r["p50_ms"], r["p99_ms"], r["guide_x_gbp_per_s"]))
        sweep.append(r)
    ok = [r for r in sweep if r["p99_ms"] <= p99_ms]
    out = {"p99_target_ms": p99_ms, "sweep": sweep,
           "capacity_guide_x_gbp_per_s": max((r["guide_x_gbp_per_s"] for r in ok), default=None),
           "capacity_qps": max((r["qps_offered"] for r in ok), default=None)}
    if plain_requests > 0:
        reqs = workloads.serving_stream(plain_requests, levels[0], "synthetic", seed=3)
        guides, busy = 0, 0.0
        for _, groups, n in reqs:
            qs = [ctx.query(t) for _, t in groups]
            t0 = time.perf_counter()
            ctx.search(qs, casoff.MODE_PLAIN).records()
            busy += time.perf_counter() - t0
            guides += n
        out["plain"] = {"requests": plain_requests, "guides": guides, "busy_s": round(busy, 3),
                        "guide_x_gbp_per_s": round(guides * ctx.info()[0] / 1e9 / busy, 2)}
        if out["capacity_guide_x_gbp_per_s"]:
            out["capacity_over_plain"] = round(out["capacity_guide_x_gbp_per_s"] / out["plain"]["guide_x_gbp_per_s"], 1)
        log("plain serving: %.0f guide×Gbp/s over %d requests" % (out["plain"]["guide_x_gbp_per_s"], plain_requests))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mixed_pam_2000", choices=sorted(WORKLOADS))
    ap.add_argument("--extra-workloads", default="genome_3g_ngg,bulge_3g_nrg",
                    help="comma-separated workloads reported beside the main one ('' for none)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--genome-len", type=int, default=3_100_000_000)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain path (profiling runs)")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=20.0)
    ap.add_argument("--serve-sweep", default="20,40,80,120,160,200,240,280,320",
                    help="offered request rates of the served-capacity sweep ('' to skip)")
    ap.add_argument("--serve-p99-ms", type=float, default=100.0, help="p99 latency bound of the served capacity")
    ap.add_argument("--serve-plain", type=int, default=12, help="requests the plain path serves beside the sweep")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the search would shard by genome "
              "range or by query, with no exchange step between GPUs).",
              file=sys.stderr)
        sys.exit(2)

    ctx = casoff.Context(0)
    t0 = time.perf_counter()
    ctx.synth(workloads.genome_3g(total_len=args.genome_len))
    t_load = time.perf_counter() - t0
    total, nfiles, ncontigs = ctx.info()
    up_ms = ctx.stats()["t_upload_ms"]
    up_bytes = 3 * 4 * ((total + 31) // 32)
    log("genome: %d positions, %d contigs, generated+uploaded in %.1f s (upload %.1f ms, %.1f GB/s)"
        % (total, ncontigs, t_load, up_ms, up_bytes / up_ms / 1e6))
    texts, main_res = run_workload(ctx, args.workload, args.steps, args.warmup, args.plain_steps, not args.no_plain)
    used = {rec["sig"] for rec in main_res.get("launches", [])}
    main_res["autotune"]["variants"] = {k: v for k, v in ctx.tuner_info().items() if k in used}
    out = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "guide×Gbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": args.workload + ": " + WORKLOADS[args.workload], "guides": main_res["guides"],
                   "queries": main_res["queries"], "genome_bp": total,
                   "mode": "batched (fused finder+comparer, one launch per pattern group)"},
        "roofline": main_res.pop("roofline"),
    }
    for k in ("plain", "batched_over_plain"):
        if k in main_res:
            out[k] = main_res.pop(k)
    for k in ("value", "ms_per_step", "guides", "queries"):
        main_res.pop(k)
    out.update(main_res)
    out["genome_generate_upload_s"] = round(t_load, 2)
    out["genome_upload"] = {"ms": round(up_ms, 1), "plane_bytes": up_bytes, "GB_per_s": round(up_bytes / up_ms / 1e6, 1),
                            "path": "pageable planes -> 2 x 64 MiB pinned staging -> HBM (hipMemcpyAsync)"}
    extra = [w for w in args.extra_workloads.split(",") if w and w != args.workload]
    if extra:
        out["workloads"] = {}
        for w in extra:
            _, res = run_workload(ctx, w, args.steps, args.warmup, max(args.plain_steps, 1), not args.no_plain)
            res.pop("launches")
            out["workloads"][w] = res
    if args.serve_requests > 0:
        log("serving: %d requests at %.1f QPS" % (args.serve_requests, args.serve_qps))
        out["serving"] = serve_bench(ctx, args.serve_requests, args.serve_qps)
        log("serving: p50 %.1f ms, p99 %.1f ms" % (out["serving"]["p50_ms"], out["serving"]["p99_ms"]))
        levels = [float(x) for x in args.serve_sweep.split(",") if x]
        if levels:
            out["serving_capacity"] = serve_capacity(ctx, args.serve_requests, levels, args.serve_p99_ms,
                                                     args.serve_plain)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ctx, texts)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
