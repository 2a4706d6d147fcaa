"""The C-ABI library loads without a GPU and exports every function that
include/casoff.h declares; host-only contexts work and search refuses them."""
import ctypes
import os
import re

import pytest

import casoff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as f:
                src = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(cof_[a-z_0-9]+)\s*\(", src, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    L = ctypes.CDLL(casoff.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_context_and_errors(tmp_path):
    c = casoff.Context(-1)
    assert casoff.lib().cof_version().decode().startswith("3.0.0")
    with pytest.raises(casoff.CofError):
        c.query("x\nNNNRG\nACGT 1\n")  # length mismatch
    q = c.query("x\nNNNRG\nACNNN 1\n")
    with pytest.raises(casoff.CofError):
        c.search([q])  # no genome
    gp = str(tmp_path / "g.fa")
    with open(gp, "w") as f:
        f.write(">a\nACGTACGAGG\n")
    c.load(gp)
    with pytest.raises(casoff.CofError, match="Host-only"):
        c.search([q])


def test_pattern_too_long_is_rejected():
    with pytest.raises(casoff.CofError, match="not supported"):
        casoff.Query(None, "x\n%s\n%s 1\n" % ("N" * 65, "A" * 65))
