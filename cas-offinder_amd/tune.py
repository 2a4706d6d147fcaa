#!/usr/bin/env python3
"""Builds the tuner file (tuned_gfx950.json beside libcasoff.so) on an MI355X.

A context times every kernel variant of a launch signature the first time it
meets it (cof_tuner_info lists the decisions).  This script meets the
signatures of the benchmark workloads (100 NGG guides, 20 NRG guides with DNA
bulge 2 / RNA bulge 1, 2,000 mixed-PAM guides in 16 groups) and of the
serving stream (requests of 1-256 guides, alone and merged as the server
merges them) on the 3.1 Gbp synthetic genome, in `--runs` fresh contexts that
start from no file, and writes, per signature, the variant most runs chose
(ties: the lowest).  Contexts opened later load the file and run no trials
for those signatures; lanes that may not run trials (serving) take the
decision of the nearest signature in it.

    python3 cas-offinder_amd/tune.py --out cas-offinder_amd/tuned_gfx950.json
"""
import argparse
import collections
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
os.environ["COF_TUNER_FILE"] = "none"  # every run starts from no decisions

import casoff  # noqa: E402
import workloads  # noqa: E402


def query_sets(ctx, n_requests, merge):
    """Query lists whose batched searches meet the signatures to tune."""
    sets = [[ctx.query(workloads.genome_3g_ngg("synthetic"))],
            [ctx.query(workloads.bulge_3g_nrg("synthetic"))],
            [ctx.query(t) for _, t in workloads.mixed_pam_2000("synthetic")]]
    reqs = [[ctx.query(t) for _, t in groups] for _, groups, _ in workloads.serving_stream(n_requests, 20.0, "synthetic")]
    sets += reqs
    for m in merge:  # consecutive requests merged into one batch, as the server does
        for i in range(0, len(reqs) - m + 1, m):
            qs = [q for r in reqs[i:i + m] for q in r]
            if len(qs) <= 64:
                sets.append(qs)
    return sets


def one_run(n_requests, merge):
    ctx = casoff.Context(0)
    try:
        ctx.synth(workloads.genome_3g())
        t0 = time.perf_counter()
        sets = query_sets(ctx, n_requests, merge)
        for qs in sets:
            ctx.search(qs, casoff.MODE_BATCHED)
        print("run: %d searches, %d signatures, %.1f s" % (len(sets), len(ctx.tuner_info()), time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        return ctx.tuner_info()
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuned_gfx950.json"))
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--merge", default="2,4,8", help="request counts merged into one batch")
    args = ap.parse_args()
    merge = [int(x) for x in args.merge.split(",") if x]
    votes = collections.defaultdict(collections.Counter)
    for _ in range(args.runs):
        for sig, v in one_run(args.requests, merge).items():
            votes[sig][v] += 1
    best = {sig: min(c.items(), key=lambda kv: (-kv[1], kv[0]))[0] for sig, c in sorted(votes.items())}
    with open(args.out, "w") as f:
        json.dump(best, f, indent=1, sort_keys=True)
        f.write("\n")
    split = sum(1 for c in votes.values() if len(c) > 1)
    print("%d signatures written to %s (%d with runs disagreeing)" % (len(best), args.out, split), file=sys.stderr)


if __name__ == "__main__":
    main()
