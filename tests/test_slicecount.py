"""The bit-sliced counting network of the bulge slice filter
(cas-offinder_amd/csrc/slicecount.hpp), compiled for the host and checked
against per-lane popcounts and comparisons (tests/cpp/slicecount_test.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_slice_count_matches_popcount(tmp_path):
    exe = str(tmp_path / "slicecount_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "cpp", "slicecount_test.cpp")],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
