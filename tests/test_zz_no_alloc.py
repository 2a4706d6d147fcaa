"""cof_search_noalloc, the batched search that allocates nothing, against
cof_search: on two calls in a row (the second with fewer queries) the hits it
writes unsorted into caller-owned device buffers equal, as a multiset, the
records of cof_search of the same queries (global key index, window start,
strand, mismatches); a buffer that is too small is refused with the number of
hits found.  The caller's device buffers come from hipMalloc of the HIP
runtime the library itself links (ctypes)."""
import ctypes

import numpy as np
import pytest

import casoff
import workloads

pytestmark = pytest.mark.gpu

COF_E_NOMEM = -2


def _sorted(keys, mm):
    order = np.lexsort((mm, keys))
    return keys[order], mm[order]


def _want(ctx, qs):
    rec = ctx.search(qs, casoff.MODE_BATCHED).records()
    base = np.cumsum([0] + [q.info()[2] for q in qs])[:-1].astype(np.uint64)
    keys = ((base[rec["query"]] + rec["key"].astype(np.uint64)) << np.uint64(40)) | \
        (rec["pos"].astype(np.uint64) << np.uint64(1)) | rec["strand"].astype(np.uint64)
    return _sorted(keys, rec["mm"].astype(np.uint16))


class DevBuf:
    """hipMalloc-ed device memory of the library's own HIP runtime."""
    hip = None

    def __init__(self, nbytes):
        if DevBuf.hip is None:
            DevBuf.hip = ctypes.CDLL("libamdhip64.so.7")
        self.p = ctypes.c_void_p()
        assert DevBuf.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def read(self, dtype, n):
        out = np.empty(n, dtype=dtype)
        if n:
            assert DevBuf.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p,
                                        ctypes.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        DevBuf.hip.hipFree(self.p)


def test_noalloc_matches_search(tmp_path):
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=25, n_contigs=4, total_len=3_000_000, n_frac=0.01), gp, 0)
    texts = [t for _, t in workloads.group_inputs(workloads.mixed_guides(7, 120, 3, 6, 0.4, 1), gp)]
    assert any(len(t.split("\n")[1].split()) == 3 for t in texts)  # bulge groups among them
    ctx = casoff.Context(0)
    try:
        ctx.load(gp)
        qs = [ctx.query(t) for t in texts]
        cap = 1 << 21
        keys, mm = DevBuf(8 * cap), DevBuf(2 * cap)
        nested = [q for q, t in zip(qs, texts) if "NGG" in t.split("\n")[1]][:3]
        for sel in (qs, nested):  # the second call has fewer queries and hits
            want = _want(ctx, sel)
            assert len(want[0]) > 0
            rc, n = ctx.search_noalloc(sel, keys.p.value, mm.p.value, cap)
            assert rc == 0, ctx.error()
            assert n == len(want[0])
            got = _sorted(keys.read(np.uint64, n), mm.read(np.uint16, n))
            assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
        # too small a buffer: refused, the count reported
        want = _want(ctx, qs)
        rc, n = ctx.search_noalloc(qs, keys.p.value, mm.p.value, len(want[0]) - 1)
        assert rc == COF_E_NOMEM and n == len(want[0])
        rc, n = ctx.search_noalloc(qs, keys.p.value, mm.p.value, len(want[0]))
        assert rc == 0 and n == len(want[0])
        keys.free()
        mm.free()
    finally:
        ctx.close()
