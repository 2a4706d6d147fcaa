#!/usr/bin/env python3
"""Benchmark of the Cas-OFFinder search path on one MI355X.

A step is one full search of the workload over the genome resident in HBM:
key tables submitted -> fused finder+comparer kernels -> compact records in
host memory (sorted).  Genome generation/upload and text formatting are not
in the step and are reported separately.

Default workload (genome_3g_ngg): synthetic 3.1 Gbp genome (seed 1, 24
contigs, GC 41%, ~5% N), pattern 20N+NGG, 100 random 20-nt guides, mm <= 4.
Prints one JSON line.
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
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s: 256 CU x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0
# The batched kernels are bound by VALU issue (integer bit work; their HBM
# reads are 0.375 B per genome position per launch).  roofline.achieved is
# the dominant kernel's executed int32 lane-ops per launch (rocprofv3
# SQ_INSTS_VALU x 64, from the committed summary of the same kernel and
# workload) over its launch time measured here.  Beside it:
#  - hbm: the algorithmic bytes (the three genome bit planes) per launch;
#  - comparison_model (counted by the library as n_model_ops): fused_kernel
#    (queries without bulges) 3 int32 ops per (candidate, compare key,
#    32-position word) (2x v_bitop3 + v_bcnt; 6 for keys with IUPAC codes);
#    bulge_kernel per (candidate, guide, word) 2 ops per window alignment (5
#    with IUPAC codes), 1 for the no-bulge popcount and 2 (AND + popcount) per
#    bulge family, plus one OR per family.  The seed-index kernels skip most
#    of that work, so their frac in this unit exceeds 1;
#  - onehot_cost_model: a 4-bit one-hot comparer (8 bases per 32-bit word,
#    about 6 ops per word), i.e. 0.75 ops per base comparison of every
#    expanded key with no early exit.
ONEHOT_OPS_PER_BASE_CMP = 0.75


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name, ctx, n_guides=None):
    if name == "genome_3g_ngg":
        texts = [workloads.genome_3g_ngg("synthetic", n_guides or 100)]
        desc = "3.1 Gbp synthetic genome (seed 1, 24 contigs), pattern 20N+NGG, %d guides, mm<=4" % (n_guides or 100)
    elif name == "bulge_3g_nrg":
        texts = [workloads.bulge_3g_nrg("synthetic", n_guides or 20)]
        desc = "3.1 Gbp synthetic genome (seed 1), 20N+NRG, DNA bulge 2 / RNA bulge 1, %d guides, mm<=4" % (n_guides or 20)
    elif name == "mixed_pam_2000":
        texts = [t for _, t in workloads.mixed_pam_2000("synthetic", n_guides or 2000)]
        desc = "3.1 Gbp synthetic genome (seed 1), %d mixed-PAM guides (seed 2), one query per (PAM, D, R) group" % (n_guides or 2000)
    else:
        raise SystemExit("unknown workload " + name)
    qs = [ctx.query(t) for t in texts]
    guides = sum(len([l for l in t.split("\n")[2:] if l.strip()]) for t in texts)
    return texts, qs, guides, desc


def kernels_of(variants):
    """Kernels the batched path launched, from the autotuner's decisions
    (signature -> variant): fused_kernel, fused_seed_kernel (variant 2) or
    fused_mask_kernel (variants 3, 4) for key groups, bulge_kernel (variants
    0, 1) or bulge_index_kernel (variant 2) for groups with DNA/RNA bulges."""
    names = set()
    for sig, v in variants.items():
        if sig.startswith("bulge|"):
            names.add("bulge_index_kernel" if v == 2 else "bulge_kernel")
        else:
            names.add({2: "fused_seed_kernel", 3: "fused_mask_kernel", 4: "fused_mask_kernel"}.get(v, "fused_kernel"))
    return "+".join(sorted(names))


def profile_record(workload, kind):
    """A figure per full-genome launch from a committed rocprofv3 summary of
    this workload (profiles/*.md, an HTML comment `<!-- kind {json} -->`):
    kind "traffic" (FETCH_SIZE / WRITE_SIZE bytes) or "valu" (SQ_INSTS_VALU
    wave-instructions of the dominant kernel)."""
    import glob
    import re
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        m = re.search(r"<!-- %s (\{.*\}) -->" % kind, open(p).read())
        if m:
            d = json.loads(m.group(1))
            if d.get("workload") == workload:
                return d, os.path.relpath(p, ROOT)
    return None, None


def measured_traffic(workload):
    d, src = profile_record(workload, "traffic")
    return (d["fetch_bytes"] + d["write_bytes"], src) if d else (None, None)


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
    tune_ms = 0.0
    for _ in range(warmup):  # the first batched search runs the autotuner's trials
        ctx.search(qs, mode)
        tune_ms += ctx.stats()["t_tune_ms"]
    time_steps.tune_ms = tune_ms
    kernel_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.search(qs, mode)
        st = ctx.stats()
        kernel_ms += st["t_fused_ms"] if mode == casoff.MODE_BATCHED else st["t_finder_ms"] + st["t_comparer_ms"]
        launches += st["n_launches"]
    t1 = time.perf_counter()
    return (t1 - t0) / steps, kernel_ms / steps, launches / steps, st, r


def cpu_baseline(ctx, texts, target_s=12.0):
    """The oracle (CPU restatement, OpenMP) on a prefix of the first contig."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    name, pos, clen, _ = ctx.contigs()[0]
    sample = min(clen, 4_000_000)
    with tempfile.TemporaryDirectory() as td:
        res = None
        for _ in range(3):
            gp = os.path.join(td, "sample.fa")
            with open(gp, "w") as f:
                f.write(">" + name + "\n")
                seq = ctx.text(pos, sample)
                for i in range(0, len(seq), 60):
                    f.write(seq[i:i + 60] + "\n")
            ip = os.path.join(td, "in.txt")
            with open(ip, "w") as f:
                f.write(gp + "\n" + texts[0].split("\n", 1)[1])
            res = oracle.time_search(ip, threads)
            if res["seconds"] >= target_s / 4 or sample >= clen:
                break
            sample = min(clen, int(sample * max(2.0, target_s / max(res["seconds"], 1e-3))))
    n_guides = len([l for l in texts[0].split("\n")[2:] if l.strip()])
    value = n_guides * (res["genome_len"] / 1e9) / res["seconds"]
    return {"value": value, "unit": "guide×Gbp/s", "cores": threads, "kind": "port",
            "sample": "oracle (reference-algorithm CPU restatement, OpenMP) on the first %d bp of %s with the "
                      "first query's %d guides: %.2f s, %d candidates" % (sample, name, n_guides, res["seconds"],
                                                                          res["candidates"])}


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
    srv.stop()
    lat = sorted(ms for _, ms in done)
    pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
    out = {"requests": n_requests, "qps_offered": qps, "guides": n_guides,
           "p50_ms": round(pct(0.50), 2), "p99_ms": round(pct(0.99), 2), "max_ms": round(lat[-1], 2),
           "wall_s": round(wall, 3), "guide_x_gbp_per_s": round(n_guides * ctx.info()[0] / 1e9 / wall, 2)}
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome_3g_ngg")
    ap.add_argument("--guides", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--genome-len", type=int, default=3_100_000_000)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain path (profiling runs)")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=20.0)
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
    log("genome: %d positions, %d contigs, generated+uploaded in %.1f s" % (total, ncontigs, t_load))
    texts, qs, n_guides, desc = build_workload(args.workload, ctx, args.guides)

    log("workload: %s (%d queries, %d guides)" % (args.workload, len(qs), n_guides))
    s_b, k_b, l_b, st_b, r_b = time_steps(ctx, qs, casoff.MODE_BATCHED, args.steps, args.warmup)
    hits_b = r_b.records()
    tune = {"trial_ms_in_warmup": round(time_steps.tune_ms, 2), "variants": ctx.tuner_info(),
            "timed_steps_launch": {0: "eager launches", 1: "captured HIP graph", 2: "replayed HIP graph"}[
                int(st_b["graph"])]}
    log("batched: %.2f ms/step (kernels %.2f ms), %d hits" % (s_b * 1e3, k_b, len(hits_b)))
    if args.no_plain:
        s_p, k_p, l_p, same = float("nan"), float("nan"), 0, None
    else:
        s_p, k_p, l_p, st_p, r_p = time_steps(ctx, qs, casoff.MODE_PLAIN, args.plain_steps, 0)
        hits_p = r_p.records()
        log("plain: %.2f ms/step (kernels %.2f ms), %d hits" % (s_p * 1e3, k_p, len(hits_p)))
        same = bool(len(hits_b) == len(hits_p) and (hits_b == hits_p).all())

    gbp = total / 1e9
    value = n_guides * gbp / s_b
    plain_value = n_guides * gbp / s_p if not args.no_plain else float("nan")
    avg_launch_ms = k_b / max(l_b, 1)
    launch_s = avg_launch_ms * 1e-3
    model = st_b["n_model_ops"] / max(l_b, 1) / launch_s / 1e12
    onehot = ONEHOT_OPS_PER_BASE_CMP * st_b["n_base_cmp"] / max(l_b, 1) / launch_s / 1e12
    hbm_achieved = 0.375 * total / launch_s / 1e9
    traffic, traffic_src = measured_traffic(args.workload)
    kernel = kernels_of(tune["variants"])
    # VALU: the lane-ops the dominant kernel executes per launch (rocprofv3
    # SQ_INSTS_VALU x 64 lanes, committed summary of the same kernel and
    # workload) over the launch time measured here; without such a summary,
    # the comparison model's ops
    vrec, vsrc = profile_record(args.workload, "valu")
    if vrec and vrec.get("kernel") == kernel and l_b == 1:
        valu_ops = vrec["valu_insts"] * 64
        work = "%d int32 lane-ops executed per launch (rocprofv3 SQ_INSTS_VALU x 64, %s)" % (valu_ops, vsrc)
    else:
        valu_ops, vsrc = st_b["n_model_ops"] / max(l_b, 1), None
        work = "%d int32 ops (%s)" % (valu_ops, {
            "fused_kernel": "3 per (candidate, key, 32-position word)",
            "bulge_kernel": "per (candidate, guide, word): 2 per alignment + 1 + 3 per bulge family",
            "bulge_index_kernel": "bulge_kernel's family model, brute-force-equivalent",
        }.get(kernel, "comparison-model units of the kernels summed"))
    achieved = valu_ops / launch_s / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "guide×Gbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_b * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": args.workload + ": " + desc, "guides": n_guides, "genome_bp": total,
                   "mode": "batched (fused finder+comparer, one launch per pattern group)"},
        "plain": None if args.no_plain else {
            "value": round(plain_value, 2), "ms_per_step": round(s_p * 1e3, 3), "steps": args.plain_steps,
            "launches_per_step": l_p, "bit_exact_vs_batched": same},
        "batched_over_plain": None if args.no_plain else round(value / plain_value, 2),
        "candidate_sites_per_s": st_b["n_pairs"] / s_b,
        "counters": {"candidates": st_b["n_candidates"], "pairs": st_b["n_pairs"], "hits": len(hits_b),
                     "base_comparisons": st_b["n_base_cmp"]},
        "genome_generate_upload_s": round(t_load, 2),
        "autotune": tune,
        "roofline": {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3),
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": traffic, "traffic_source": traffic_src, "avg_launch_ms": round(avg_launch_ms, 4),
                     "work_per_launch": work,
                     "hbm": {"algorithmic_bytes_per_launch": int(0.375 * total), "achieved": round(hbm_achieved, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(hbm_achieved / HBM_PEAK_GBS, 4)},
                     "comparison_model": {
                         "ops_per_launch": int(st_b["n_model_ops"] / max(l_b, 1)), "achieved": round(model, 3),
                         "frac": round(model / VALU_PEAK_TOPS, 4),
                         "unit": "3 int32 ops per (candidate, key, 32-position word) for key tables, the bulge "
                                 "kernel's family model for bulge tables: brute-force-equivalent work, which the "
                                 "seed indexes skip (frac above 1)"},
                     "onehot_cost_model": {"ops_per_base_cmp": ONEHOT_OPS_PER_BASE_CMP, "achieved": round(onehot, 2),
                                           "frac": round(onehot / VALU_PEAK_TOPS, 3)}},
    }
    if not args.no_plain and st_p["t_finder_ms"] > 0:
        # plain path's finder (finder_count_kernel + tile-count scan +
        # finder_write_kernel): algorithmic bytes = 0.375 B read per scanned
        # position + a (locus u32, flag u8) entry written per locus with a PAM
        fbytes = 0.375 * st_p["n_finder_positions"] + 5 * st_p["n_loci"]
        fach = fbytes / (st_p["t_finder_ms"] * 1e-3) / 1e9
        out["finder_roofline"] = {"kernel": "finder_count_kernel + finder_write_kernel (plain path)", "bound": "hbm",
                                  "algorithmic_bytes_per_step": int(fbytes), "ms_per_step": round(st_p["t_finder_ms"], 4),
                                  "achieved": round(fach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(fach / HBM_PEAK_GBS, 4)}
    if args.serve_requests > 0:
        log("serving: %d requests at %.1f QPS" % (args.serve_requests, args.serve_qps))
        out["serving"] = serve_bench(ctx, args.serve_requests, args.serve_qps)
        log("serving: p50 %.1f ms, p99 %.1f ms" % (out["serving"]["p50_ms"], out["serving"]["p99_ms"]))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ctx, texts)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
