"""ctypes binding of libcasoff.so (include/casoff.h) for tests, bench and Python users.

The library is the product: HIP kernels for gfx950 behind a C ABI.  This module
only marshals arguments; loading fails loudly when the library is not built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("COF_LIB") or os.path.join(HERE, "libcasoff.so")
CLI_PATH = os.path.join(HERE, "cas-offinder")

MODE_PLAIN = 0
MODE_BATCHED = 1

HIT_DTYPE = np.dtype([("pos", "<u8"), ("key", "<u4"), ("mm", "<u2"), ("strand", "u1"), ("query", "u1")])


class SynthSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n_contigs", C.c_uint32),
        ("length_mode", C.c_uint32),
        ("total_len", C.c_uint64),
        ("min_len", C.c_uint64),
        ("max_len", C.c_uint64),
        ("gc", C.c_double),
        ("n_frac", C.c_double),
        ("n_run_min", C.c_uint32),
        ("n_run_max", C.c_uint32),
        ("lower_frac", C.c_double),
        ("line_width", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("t_total_ms", C.c_double),
        ("t_finder_ms", C.c_double),
        ("t_comparer_ms", C.c_double),
        ("t_fused_ms", C.c_double),
        ("n_candidates", C.c_uint64),
        ("n_pairs", C.c_uint64),
        ("n_hits", C.c_uint64),
        ("n_launches", C.c_uint64),
        ("genome_len", C.c_uint64),
        ("n_base_cmp", C.c_uint64),
        ("n_model_ops", C.c_uint64),
        ("t_collect_ms", C.c_double),
        ("n_loci", C.c_uint64),
        ("n_finder_positions", C.c_uint64),
        ("t_tune_ms", C.c_double),
        ("graph", C.c_uint64),
        ("t_sort_ms", C.c_double),
        ("t_copy_ms", C.c_double),
        ("t_decode_ms", C.c_double),
        ("t_upload_ms", C.c_double),
        ("t_plan_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerCfg(C.Structure):
    _fields_ = [("mem_budget_bytes", C.c_uint64), ("max_batch_queries", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("lane_hits", C.c_uint64), ("batches", C.c_uint64), ("requests", C.c_uint64),
                ("queries", C.c_uint64), ("hits", C.c_uint64), ("lane_grows", C.c_uint64),
                ("admission_waits", C.c_uint64), ("peak_inflight_bytes", C.c_uint64),
                ("hits_per_key", C.c_double), ("launches", C.c_uint64), ("t_search_ms", C.c_double),
                ("t_plan_ms", C.c_double), ("t_kernel_ms", C.c_double), ("t_split_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CofError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CofError("libcasoff.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, u64, sz = C.c_void_p, C.c_int, C.c_uint64, C.c_size_t
        sig = {
            "cof_version": (C.c_char_p, []),
            "cof_open": (i32, [i32, C.POINTER(vp)]),
            "cof_device_ids": (i32, [C.c_char_p, i32, C.POINTER(C.c_uint8), C.POINTER(i32)]),
            "cof_device_count": (i32, [C.POINTER(i32)]),
            "cof_close": (i32, [vp]),
            "cof_last_error": (C.c_char_p, [vp]),
            "cof_genome_load": (i32, [vp, C.c_char_p, C.POINTER(i32)]),
            "cof_genome_synth": (i32, [vp, C.POINTER(SynthSpec)]),
            "cof_genome_info": (i32, [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
            "cof_genome_contig": (i32, [vp, u64, C.POINTER(C.c_char_p), C.POINTER(u64), C.POINTER(u64),
                                    C.POINTER(C.c_uint32)]),
            "cof_genome_text": (i32, [vp, u64, u64, C.c_char_p]),
            "cof_synth_write": (i32, [C.POINTER(SynthSpec), C.c_char_p, i32]),
            "cof_query_parse": (i32, [vp, C.c_char_p, sz, C.POINTER(vp)]),
            "cof_query_free": (i32, [vp]),
            "cof_query_dump": (i32, [vp, C.POINTER(C.c_void_p), C.POINTER(sz)]),
            "cof_query_info": (i32, [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
            "cof_search": (i32, [vp, C.POINTER(vp), i32, i32, C.POINTER(vp)]),
            "cof_stats_get": (i32, [vp, C.POINTER(Stats)]),
            "cof_search_noalloc": (i32, [vp, C.POINTER(vp), i32, vp, vp, u64, C.POINTER(u64)]),
            "cof_tuner_info": (i32, [vp, C.POINTER(C.c_void_p), C.POINTER(sz)]),
            "cof_launch_info": (i32, [vp, C.POINTER(C.c_void_p), C.POINTER(sz)]),
            "cof_tuner_load": (i32, [vp, C.c_char_p]),
            "cof_tuner_save": (i32, [vp, C.c_char_p]),
            "cof_result_count": (C.c_int64, [vp]),
            "cof_result_records": (i32, [vp, vp, sz]),
            "cof_result_format": (i32, [vp, vp, i32, C.c_char_p, i32, C.c_char_p]),
            "cof_result_format_mem": (i32, [vp, vp, i32, C.POINTER(C.c_void_p), C.POINTER(sz)]),
            "cof_result_free": (i32, [vp]),
            "cof_format_hits": (i32, [vp, vp, vp, sz, C.POINTER(C.c_void_p), C.POINTER(sz)]),
            "cof_free": (None, [vp]),
            "cof_server_start": (i32, [vp, C.POINTER(ServerCfg), C.POINTER(vp)]),
            "cof_server_submit": (i32, [vp, C.POINTER(vp), i32, C.POINTER(u64)]),
            "cof_server_poll": (i32, [vp, u64, i32, C.POINTER(vp), C.POINTER(C.c_double)]),
            "cof_server_stop": (i32, [vp]),
            "cof_server_stats_get": (i32, [vp, C.POINTER(ServerStats)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _take_buffer(ptr, n):
    try:
        return C.string_at(ptr, n).decode("latin-1")
    finally:
        lib().cof_free(ptr)


def device_ids(devarg, n_devices):
    """Device indices (< n_devices) that the CLI's device-id text selects, as
    Cas_OFFinder(devtype, devarg) reads it (cas-offinder.cpp:131-153), and the
    number of malformed pieces."""
    sel = (C.c_uint8 * max(n_devices, 1))()
    bad = C.c_int(0)
    rc = lib().cof_device_ids(devarg.encode(), n_devices, sel, C.byref(bad))
    if rc:
        raise CofError("cof_device_ids(%r) failed (%d)" % (devarg, rc))
    return [d for d in range(n_devices) if sel[d]], bad.value


def synth_spec(**kw):
    s = SynthSpec()
    defaults = dict(seed=0, n_contigs=1, length_mode=0, total_len=1000, min_len=0, max_len=0, gc=0.41,
                    n_frac=0.0, n_run_min=100, n_run_max=5000, lower_frac=0.0, line_width=60)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(s, k, v)
    return s


def synth_write(spec, path, fmt=0):
    rc = lib().cof_synth_write(C.byref(spec), path.encode(), fmt)
    if rc:
        raise CofError("cof_synth_write failed: %s" % lib().cof_last_error(None).decode())


class Query:
    def __init__(self, ctx, text):
        if isinstance(text, str):
            text = text.encode("latin-1")
        self._p = C.c_void_p()
        rc = lib().cof_query_parse(ctx._p if ctx else None, text, len(text), C.byref(self._p))
        if rc:
            msg = lib().cof_last_error(ctx._p if ctx else None).decode()
            raise CofError(msg)

    def dump(self):
        out, ln = C.c_void_p(), C.c_size_t()
        if lib().cof_query_dump(self._p, C.byref(out), C.byref(ln)):
            raise CofError("cof_query_dump failed")
        return _take_buffer(out, ln.value)

    def info(self):
        d, p, n = C.c_char_p(), C.c_uint32(), C.c_uint32()
        lib().cof_query_info(self._p, C.byref(d), C.byref(p), C.byref(n))
        return d.value.decode("latin-1"), p.value, n.value

    def __del__(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.cof_query_free(self._p)
            self._p = C.c_void_p()


class Result:
    def __init__(self, ctx, p):
        self.ctx = ctx
        self._p = p

    def count(self):
        return lib().cof_result_count(self._p)

    def records(self):
        n = self.count()
        arr = np.zeros(n, dtype=HIT_DTYPE)
        if n:
            lib().cof_result_records(self._p, arr.ctypes.data, n)
        return arr

    def format(self, qi=0):
        out, ln = C.c_void_p(), C.c_size_t()
        rc = lib().cof_result_format_mem(self.ctx._p, self._p, qi, C.byref(out), C.byref(ln))
        if rc:
            raise CofError(self.ctx.error())
        return _take_buffer(out, ln.value)

    def write(self, qi, path, append=False, summary=""):
        rc = lib().cof_result_format(self.ctx._p, self._p, qi, path.encode(), int(append), summary.encode())
        if rc:
            raise CofError(self.ctx.error())

    def __del__(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.cof_result_free(self._p)
            self._p = C.c_void_p()


class Context:
    def __init__(self, device=0):
        self._p = C.c_void_p()
        rc = lib().cof_open(device, C.byref(self._p))
        if rc:
            raise CofError("cof_open(%d) failed: %s" % (device, lib().cof_last_error(None).decode()))

    def error(self):
        return lib().cof_last_error(self._p).decode("latin-1")

    def close(self):
        if self._p.value:
            if lib().cof_close(self._p):  # refused while a server is attached
                raise CofError(self.error())
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load(self, path):
        n = C.c_int()
        rc = lib().cof_genome_load(self._p, path.encode(), C.byref(n))
        if rc:
            raise CofError(self.error())
        return n.value

    def synth(self, spec):
        rc = lib().cof_genome_synth(self._p, C.byref(spec))
        if rc:
            raise CofError(self.error())

    def info(self):
        t, f, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib().cof_genome_info(self._p, C.byref(t), C.byref(f), C.byref(c))
        return t.value, f.value, c.value

    def contigs(self):
        out = []
        i = 0
        while True:
            nm, p, ln, f = C.c_char_p(), C.c_uint64(), C.c_uint64(), C.c_uint32()
            if lib().cof_genome_contig(self._p, i, C.byref(nm), C.byref(p), C.byref(ln), C.byref(f)):
                return out
            out.append((nm.value.decode("latin-1"), p.value, ln.value, f.value))
            i += 1

    def text(self, pos, n):
        buf = C.create_string_buffer(n)
        rc = lib().cof_genome_text(self._p, pos, n, buf)
        if rc:
            raise CofError("cof_genome_text failed")
        return buf.raw.decode("latin-1")

    def query(self, text):
        return Query(self, text)

    def search(self, queries, mode=MODE_BATCHED):
        arr = (C.c_void_p * len(queries))(*[q._p.value for q in queries])
        out = C.c_void_p()
        rc = lib().cof_search(self._p, arr, len(queries), mode, C.byref(out))
        if rc:
            raise CofError(self.error())
        return Result(self, out)

    def search_noalloc(self, queries, hit_keys_ptr, hit_mm_ptr, cap):
        """cof_search_noalloc into caller-owned device buffers (raw device
        pointers); returns (status, hits found)."""
        arr = (C.c_void_p * len(queries))(*[q._p.value for q in queries])
        n = C.c_uint64()
        rc = lib().cof_search_noalloc(self._p, arr, len(queries), hit_keys_ptr, hit_mm_ptr, cap, C.byref(n))
        return rc, n.value

    def stats(self):
        s = Stats()
        lib().cof_stats_get(self._p, C.byref(s))
        return s.as_dict()

    def tuner_info(self):
        """{launch signature: chosen kernel variant} of the autotuner so far."""
        out, ln = C.c_void_p(), C.c_size_t()
        if lib().cof_tuner_info(self._p, C.byref(out), C.byref(ln)):
            raise CofError(self.error())
        text = _take_buffer(out, ln.value)
        return dict((l.rsplit("\t", 1)[0], int(l.rsplit("\t", 1)[1])) for l in text.splitlines() if l)

    def tuner_load(self, path):
        if lib().cof_tuner_load(self._p, path.encode()):
            raise CofError(self.error())

    def tuner_save(self, path):
        if lib().cof_tuner_save(self._p, path.encode()):
            raise CofError(self.error())

    def launch_info(self):
        """Batched launches of the last search: [{kernel, sig, variant, ms, candidates}]."""
        out, ln = C.c_void_p(), C.c_size_t()
        if lib().cof_launch_info(self._p, C.byref(out), C.byref(ln)):
            raise CofError(self.error())
        recs = []
        for l in _take_buffer(out, ln.value).splitlines():
            k, sig, v, ms, cand = l.split("\t")
            recs.append({"kernel": k, "sig": sig, "variant": int(v), "ms": float(ms), "candidates": int(cand)})
        return recs

    def format_hits(self, query, hits):
        hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
        out, ln = C.c_void_p(), C.c_size_t()
        rc = lib().cof_format_hits(self._p, query._p, hits.ctypes.data if len(hits) else None, len(hits),
                                   C.byref(out), C.byref(ln))
        if rc:
            raise CofError(self.error())
        return _take_buffer(out, ln.value)

    def __del__(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            try:
                self.close()
            except CofError:
                pass


class Server:
    """Continuous-batching request loop over one context (cof_server_*)."""

    def __init__(self, ctx, mem_budget_bytes=8 << 30, max_batch_queries=64):
        self.ctx = ctx
        cfg = ServerCfg(mem_budget_bytes, max_batch_queries)
        self._p = C.c_void_p()
        if lib().cof_server_start(ctx._p, C.byref(cfg), C.byref(self._p)):
            raise CofError(ctx.error())

    def submit(self, queries):
        arr = (C.c_void_p * len(queries))(*[q._p.value for q in queries])
        t = C.c_uint64()
        if lib().cof_server_submit(self._p, arr, len(queries), C.byref(t)):
            raise CofError("submit failed")
        return t.value

    def poll(self, ticket, wait=True):
        """(Result, latency_ms) when done, None while pending."""
        out, lat = C.c_void_p(), C.c_double()
        rc = lib().cof_server_poll(self._p, ticket, int(wait), C.byref(out), C.byref(lat))
        if rc == 1:
            return None
        if rc:
            raise CofError(self.ctx.error())
        return Result(self.ctx, out), lat.value

    def stats(self):
        st = ServerStats()
        if lib().cof_server_stats_get(self._p, C.byref(st)):
            raise CofError("server stats failed")
        return st.as_dict()

    def stop(self):
        if self._p.value:
            lib().cof_server_stop(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            self.stop()


def record_lines(text):
    """Non-header output lines, bytewise sorted (the canonical multiset form)."""
    return sorted(l for l in text.split("\n") if l and not l.startswith("#"))
