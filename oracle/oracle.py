"""TEST INFRASTRUCTURE: Python handle on the CPU oracle (liboracle.so and the
cas-offinder-oracle binary built from casoff_oracle.c).  Used only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker."""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
BIN = os.path.join(HERE, "cas-offinder-oracle")
REF_READERS = os.path.join(HERE, "_ref", "ref_readers")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(LIB)
        L.orc_parse_dump.restype = C.c_int
        L.orc_parse_dump.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_iupac_mismatch.restype = C.c_int
        L.orc_iupac_mismatch.argtypes = [C.c_char, C.c_char]
        L.orc_revcomp.argtypes = [C.c_char_p, C.c_size_t]
        L.orc_genome_dump.restype = C.c_void_p
        L.orc_genome_dump.argtypes = [C.c_char_p, C.POINTER(C.c_size_t)]
        L.orc_time_search.restype = C.c_double
        L.orc_time_search.argtypes = [C.c_char_p, C.c_int] + [C.POINTER(C.c_uint64)] * 4
        _lib = L
    return _lib


def iupac_mismatch(p, c):
    return bool(lib().orc_iupac_mismatch(p.encode(), c.encode()))


def revcomp(s):
    b = C.create_string_buffer(s.encode("latin-1"), len(s))
    lib().orc_revcomp(b, len(s))
    return b.raw.decode("latin-1")


def parse_dump(text):
    """(ok, dump-or-error) for a reference input text."""
    if isinstance(text, str):
        text = text.encode("latin-1")
    out = C.c_void_p()
    rc = lib().orc_parse_dump(text, len(text), C.byref(out))
    s = C.string_at(out).decode("latin-1")
    lib().orc_free(out)
    return rc == 0, s


def run(input_path, output_path, summary_path="", threads=0, timeout=600):
    """Runs the oracle CLI (main.cpp semantics); returns (exit code, stderr)."""
    args = [BIN, input_path, "C", output_path]
    if summary_path:
        args += ["--summary", summary_path]
    env = dict(os.environ)
    if threads:
        env["ORC_THREADS"] = str(threads)
    p = subprocess.run(args, capture_output=True, timeout=timeout, env=env)
    return p.returncode, p.stderr.decode("latin-1")


def time_search(input_path, threads=0):
    g, c, p, h = (C.c_uint64() for _ in range(4))
    t = lib().orc_time_search(input_path.encode(), threads, C.byref(g), C.byref(c), C.byref(p), C.byref(h))
    return dict(seconds=t, genome_len=g.value, candidates=c.value, pairs=p.value, hits=h.value)


def ref_readers(path):
    """The reference's own read_fasta/read_twobit (oracle/_ref) on one file:
    (status, [(chrpos, name)], text), or None when _ref is not built."""
    if not os.path.exists(REF_READERS):
        return None
    out = subprocess.run([REF_READERS, path], capture_output=True, timeout=120).stdout
    return _parse_reader_dump(out)


def _parse_reader_dump(out):
    lines = out.split(b"\n")
    status = int(lines[0].split()[1])
    if status:
        return status, [], ""
    nchr = int(lines[1].split()[1])
    chrs = []
    for i in range(nchr):
        parts = lines[2 + i].split(b" ", 2)
        chrs.append((int(parts[1]), parts[2].decode("latin-1") if len(parts) > 2 else ""))
    ln = int(lines[2 + nchr].split()[1])
    head = b"\n".join(lines[: 3 + nchr]) + b"\n"
    return 0, chrs, out[len(head): len(head) + ln].decode("latin-1")


def genome_dump(path):
    """The oracle's readers on one file: (status, [(chrpos, name)], text)."""
    n = C.c_size_t()
    p = lib().orc_genome_dump(path.encode(), C.byref(n))
    out = C.string_at(p, n.value)
    lib().orc_free(p)
    return _parse_reader_dump(out)
