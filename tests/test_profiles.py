"""The rocprofv3 summaries under profiles/ that bench.py reads: every record
parses, names a workload bench.py runs and a kernel it reports, holds
positive per-launch figures, and each summary lists the full command of
every run it quotes."""
import glob
import importlib.util
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORKLOADS = {"genome_3g_ngg", "bulge_3g_nrg", "mixed_pam_2000"}
KERNELS = {"fused_kernel", "fused_seed_kernel", "fused_mask_kernel", "bulge_kernel", "bulge_index_kernel",
           "bulge_slice_kernel"}


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _records():
    out = []
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for m in re.finditer(r"<!-- kernel (\{.*\}) -->", open(p).read()):
            out.append((p, json.loads(m.group(1))))
    return out


def test_kernel_records_and_bench_lookup():
    """Every per-kernel record names a workload bench.py runs and a kernel
    instance of the library, carries positive per-launch figures (a
    full-genome launch reads at least the 1.1625 GB of bit planes once), and
    is what bench.py finds for that (workload, kernel)."""
    bench = _bench()
    recs = _records()
    assert "mixed_pam_2000" in {r["workload"] for _, r in recs}
    for p, r in recs:
        assert r["workload"] in WORKLOADS, p
        assert r["kernel"].startswith("cof::") and r["kernel"].split("<")[0][5:] in KERNELS, p
        assert r["avg_us"] > 0 and r["valu_insts"] > 0, p
        assert r["fetch_bytes"] >= 1.1625e9 and r["write_bytes"] > 0, p
        got, src = bench.profile_kernel(r["workload"], r["kernel"])
        assert got == r and src == os.path.relpath(p, ROOT)
    assert bench.profile_kernel("no_such_workload", "cof::x") == (None, None)


def test_summaries_quote_their_commands():
    for p in glob.glob(os.path.join(ROOT, "profiles", "*.md")):
        text = open(p).read()
        for run in ("--kernel-trace --stats", "--pmc FETCH_SIZE", "--pmc WRITE_SIZE", "--calibrate-copy"):
            assert "rocprofv3 " + run in text or run in text, (p, run)
        assert re.search(r"```\n\{\"metric\"", text), p  # the bench.py line of the same build
