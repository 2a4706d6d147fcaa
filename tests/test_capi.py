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


# (device-id text, expected selection among 8 devices, malformed pieces), each
# worked out by hand from Cas_OFFinder's constructor (cas-offinder.cpp:131-153:
# getline splits on ',' then ':', atoi, for (j = start; j < end; j += step))
# and initOpenCLDevices (:80-104: an empty id list selects every device)
DEVICE_ID_CASES = [
    ("", list(range(8)), 0),            # no id listed: every device
    ("3", [3], 0),
    ("1,5", [1, 5], 0),
    ("0:4:2", [0, 2], 0),
    ("2:5", [2, 3, 4], 0),
    (":3", [0, 1, 2], 0),               # atoi("") = 0
    ("6:", [6], 0),                     # trailing ':' makes no item: an id
    ("4,", [4], 0),                     # trailing ',' makes no item
    ("x", [0], 0),                      # atoi("x") = 0
    ("-1", [], 0),                      # listed, as an unsigned index past every device
    ("9", [], 0),
    ("3:1", list(range(8)), 0),         # empty range lists nothing: every device
    ("3:1,2", [2], 0),
    ("1:2:3:4", list(range(8)), 1),     # malformed piece skipped, nothing listed
    ("0,1:2:3:4", [0], 1),
    (",,", list(range(8)), 2),          # two empty pieces
    ("5:6:-2", [1, 3, 5], 0),           # 5, 3, 1, then the unsigned wrap ends it
    ("1:8:3", [1, 4, 7], 0),
]


@pytest.mark.parametrize("arg,want,bad", DEVICE_ID_CASES)
def test_device_id_argument(arg, want, bad):
    assert casoff.device_ids(arg, 8) == (want, bad)


def test_device_id_argument_limits():
    assert casoff.device_ids("0:4:2", 1) == ([0], 0)
    assert casoff.device_ids("", 0) == ([], 0)
    with pytest.raises(casoff.CofError):
        casoff.device_ids("0:4:0", 8)  # the reference never leaves this loop
    assert casoff.device_ids("4:4:0", 8) == (list(range(8)), 0)  # empty range: loop not entered


def test_cli_arguments_without_search(tmp_path):
    """Argument handling of main.cpp:88-124 that needs no device: usage on
    fewer than three positional arguments or an unknown --option (exit 0),
    C/A and unknown letters refused (exit 1), and a device-id list that
    selects no present device refused (exit 1)."""
    import subprocess
    run = lambda *a: subprocess.run([casoff.CLI_PATH, *a], capture_output=True, timeout=60)
    p = run("in.txt", "G")
    assert p.returncode == 0 and b"Usage:" in p.stdout
    p = run("in.txt", "G", "out.txt", "--bogus")
    assert p.returncode == 0 and b"Usage:" in p.stdout
    for dev in ("C", "A", "X0"):
        assert run("in.txt", dev, str(tmp_path / "o.txt")).returncode == 1
    p = run("in.txt", "G4096", str(tmp_path / "o.txt"))
    assert p.returncode == 1 and b"No HIP devices found." in p.stderr
    p = run("in.txt", "G0:4:0", str(tmp_path / "o.txt"))
    assert p.returncode == 1 and b"device ID" in p.stderr


def test_failed_load_leaves_the_genome(tmp_path):
    """A load that fails (missing path) keeps the loaded genome; one that
    succeeds replaces it."""
    c = casoff.Context(-1)
    ga, gb = str(tmp_path / "a.fa"), str(tmp_path / "b.fa")
    with open(ga, "w") as f:
        f.write(">a\nACGTACGAGG\n>b\nTTTT\n")
    with open(gb, "w") as f:
        f.write(">c\nGGGGGG\n")
    c.load(ga)
    info, contigs = c.info(), c.contigs()
    with pytest.raises(casoff.CofError, match="opening directory"):
        c.load(str(tmp_path / "missing"))
    assert c.info() == info and c.contigs() == contigs
    assert c.text(0, 15) == "ACGTACGAGG;TTTT"
    c.load(gb)
    assert [x[0] for x in c.contigs()] == ["c"] and c.text(0, 6) == "GGGGGG"
    c.close()
