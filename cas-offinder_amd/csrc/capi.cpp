// extern "C" entry points of libcasoff.so (declared in include/casoff.h).
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "format.hpp"

using namespace cof;

static thread_local std::string g_noctx_err;

// The tuner file a context starts from: $COF_TUNER_FILE ("" or "none": no
// file), else tuned_gfx950.json beside the library when it exists.
static std::string default_tuner_file() {
    if (const char *e = getenv("COF_TUNER_FILE")) return strcmp(e, "none") == 0 ? std::string() : std::string(e);
    Dl_info info{};
    if (!dladdr((void *)&default_tuner_file, &info) || !info.dli_fname) return std::string();
    std::string p = info.dli_fname;
    const size_t sl = p.rfind('/');
    p = (sl == std::string::npos ? std::string(".") : p.substr(0, sl)) + "/tuned_gfx950.json";
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) return std::string();
    fclose(f);
    return p;
}

extern "C" {

const char *cof_version(void) { return "3.0.0-mi355x"; }

int cof_device_count(int *n) {
    if (!n) return COF_E_INVALID;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = std::max(c, 0);
    return COF_OK;
}

int cof_device_ids(const char *devarg, int n_devices, uint8_t *selected, int *n_bad) {
    if (!devarg || n_devices < 0 || (n_devices > 0 && !selected)) return COF_E_INVALID;
    // getline-style split: no item for an empty text or a trailing delimiter
    auto split = [](const std::string &t, char d) {
        std::vector<std::string> v;
        size_t b = 0;
        while (b < t.size()) {
            size_t e = t.find(d, b);
            if (e == std::string::npos) e = t.size();
            v.push_back(t.substr(b, e - b));
            b = e + 1;
        }
        return v;
    };
    std::vector<uint8_t> sel(n_devices, 0);
    bool listed = false;  // the reference's dev_ids is non-empty
    int bad = 0;
    for (const std::string &piece : split(devarg, ',')) {
        const std::vector<std::string> ix = split(piece, ':');
        if (ix.size() == 1) {  // one id; a negative one becomes a huge unsigned index
            listed = true;
            const unsigned id = (unsigned)atoi(ix[0].c_str());
            if (id < (unsigned)n_devices) sel[id] = 1;
        } else if (ix.size() == 2 || ix.size() == 3) {  // for (j = start; j < end; j += step)
            const unsigned a = (unsigned)atoi(ix[0].c_str()), b = (unsigned)atoi(ix[1].c_str());
            const int step = ix.size() == 3 ? atoi(ix[2].c_str()) : 1;
            if (a >= b) continue;
            if (step == 0) return COF_E_INVALID;
            listed = true;
            for (int d = 0; d < n_devices; d++) {
                const unsigned u = (unsigned)d;
                if (step > 0 ? (u >= a && u < b && (u - a) % (unsigned)step == 0)
                             : (u <= a && (a - u) % (0u - (unsigned)step) == 0))
                    sel[d] = 1;
            }
        } else {
            bad++;
        }
    }
    for (int d = 0; d < n_devices; d++) selected[d] = listed ? sel[d] : 1;
    if (n_bad) *n_bad = bad;
    return COF_OK;
}

int cof_open(int device, cof_ctx **out) {
    if (!out) return COF_E_INVALID;
    *out = nullptr;
    const std::string tf = default_tuner_file();
    if (device == -1) {  // host-only context
        auto *c = new cof_ctx();
        c->device = -1;
        if (!tf.empty() && c->tuner.load(tf, g_noctx_err)) {
            delete c;
            return COF_E_INVALID;
        }
        *out = c;
        return COF_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_noctx_err = "No HIP devices found.";
        return COF_E_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_noctx_err = "Device index out of range.";
        return COF_E_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) return COF_E_DEVICE;
    auto *c = new cof_ctx();
    c->device = device;
    if (const char *at = getenv("COF_AUTOTUNE")) c->tuner.enabled = strcmp(at, "0") != 0;
    if (const char *fv = getenv("COF_FORCE_VARIANT")) c->tuner.force = std::min(std::max(atoi(fv), 0), cof::kFusedVariants - 1);
    if (const char *gr = getenv("COF_GRAPH")) c->lane.graphs = strcmp(gr, "0") != 0;
    if (const char *hp = getenv("COF_HIT_PACK")) c->lane.pack_hits = strcmp(hp, "0") != 0;
    if (!tf.empty() && c->tuner.load(tf, g_noctx_err)) {
        delete c;
        return COF_E_INVALID;
    }
    if (noalloc_init(device) != hipSuccess) {
        delete c;
        g_noctx_err = "cannot look up the scratch array's device address";
        return COF_E_DEVICE;
    }
    if (c->lane.init() != hipSuccess) {
        c->lane.release();
        delete c;
        g_noctx_err = "cannot create a HIP stream";
        return COF_E_DEVICE;
    }
    *out = c;
    return COF_OK;
}

int cof_close(cof_ctx *c) {
    if (!c) return COF_E_INVALID;
    if (c->servers.load() > 0) return c->fail(COF_E_INVALID, "A server is attached to this context.");
    if (c->device < 0) {
        delete c;
        return COF_OK;
    }
    (void)hipSetDevice(c->device);
    c->lane.release();
    c->d_lo.release(); c->d_hi.release(); c->d_ot.release(); c->d_poison.release();
    c->d_tile_ppos.release(); c->d_tile_pidx.release();
    delete c;
    return COF_OK;
}

const char *cof_last_error(const cof_ctx *c) { return c ? c->err.c_str() : g_noctx_err.c_str(); }

// A load that fails leaves the context's genome as it was; one that succeeds
// replaces it (results searched before then no longer format).
int cof_genome_load(cof_ctx *c, const char *path, int *nfiles) {
    if (!c || !path) return COF_E_INVALID;
    if (c->servers.load() > 0) return c->fail(COF_E_INVALID, "A server is attached to this context.");
    if (c->device >= 0) (void)hipSetDevice(c->device);
    int nf = 0;
    std::string err;
    const char *v = getenv("COF_VERBOSE");
    HostGenome g;
    int rc = load_path(path, g, nf, err, v && atoi(v) > 0);
    if (rc) return c->fail(rc, err);
    if (nfiles) *nfiles = nf;
    return c->install_genome(std::move(g));
}

int cof_genome_synth(cof_ctx *c, const cof_synth_spec *spec) {
    if (!c || !spec) return COF_E_INVALID;
    if (c->servers.load() > 0) return c->fail(COF_E_INVALID, "A server is attached to this context.");
    if (c->device >= 0) (void)hipSetDevice(c->device);
    HostGenome g;
    synth_genome(*spec, g);
    return c->install_genome(std::move(g));
}

int cof_genome_info(const cof_ctx *c, uint64_t *total_len, uint64_t *n_files, uint64_t *n_contigs) {
    if (!c) return COF_E_INVALID;
    if (total_len) *total_len = c->genome.total;
    if (n_files) *n_files = c->genome.files.size();
    if (n_contigs) *n_contigs = c->genome.n_contigs();
    return COF_OK;
}

int cof_genome_contig(const cof_ctx *c, uint64_t i, const char **name, uint64_t *pos, uint64_t *len,
                      uint32_t *file) {
    if (!c) return COF_E_INVALID;
    for (size_t f = 0; f < c->genome.files.size(); f++) {
        const GenomeFile &gf = c->genome.files[f];
        if (i < gf.names.size()) {
            const uint64_t b = gf.chrpos[i];
            const uint64_t e = i + 1 < gf.names.size() ? gf.chrpos[i + 1] - 1 : gf.len;
            if (name) *name = gf.names[i].c_str();
            if (pos) *pos = gf.base + b;
            if (len) *len = e >= b ? e - b : 0;
            if (file) *file = (uint32_t)f;
            return COF_OK;
        }
        i -= gf.names.size();
    }
    return COF_E_INVALID;
}

int cof_genome_text(const cof_ctx *c, uint64_t pos, uint64_t n, char *out) {
    if (!c || !out || pos + n > c->genome.total) return COF_E_INVALID;
    c->genome.text(pos, n, out);
    return COF_OK;
}

int cof_synth_write(const cof_synth_spec *spec, const char *path, int format) {
    if (!spec || !path || (format != 0 && format != 1)) return COF_E_INVALID;
    std::string err;
    int rc = synth_write(*spec, path, format, err);
    if (rc) g_noctx_err = err;
    return rc;
}

int cof_query_parse(cof_ctx *c, const char *text, size_t len, cof_query **out) {
    if (!out || (!text && len)) return COF_E_INVALID;
    *out = nullptr;
    auto q = std::make_shared<Query>();
    std::string err;
    if (parse_query(text ? text : "", len, *q, err) || encode_query(*q, err)) {
        if (c) c->err = err;
        else g_noctx_err = err;
        return COF_E_INVALID;
    }
    *out = new cof_query{q};
    return COF_OK;
}

int cof_query_free(cof_query *q) {
    delete q;
    return COF_OK;
}

int cof_query_dump(const cof_query *q, char **out, size_t *len) {
    if (!q || !out || !len) return COF_E_INVALID;
    const Query &qq = *q->q;
    std::string t = std::to_string(qq.P) + " " + std::to_string(qq.D) + " " + std::to_string(qq.R) + " " +
                    qq.pattern + "\n";
    for (auto &k : qq.keys) {
        t += k.seq + "\t" + k.id + "\t" + std::to_string(k.threshold) + "\t";
        for (size_t j = 0; j < k.infos.size(); j++) {
            if (j) t += ",";
            t += k.infos[j].first + ":" + std::to_string(k.infos[j].second);
        }
        t += "\n";
    }
    char *p = (char *)malloc(t.size() + 1);
    if (!p) return COF_E_NOMEM;
    memcpy(p, t.data(), t.size() + 1);
    *out = p;
    *len = t.size();
    return COF_OK;
}

int cof_query_info(const cof_query *q, const char **chrdir, uint32_t *pattern_len, uint32_t *n_keys) {
    if (!q) return COF_E_INVALID;
    if (chrdir) *chrdir = q->q->chrdir.c_str();
    if (pattern_len) *pattern_len = q->q->P;
    if (n_keys) *n_keys = (uint32_t)q->q->keys.size();
    return COF_OK;
}

int cof_search(cof_ctx *c, cof_query *const *queries, int n_queries, int mode, cof_result **out) {
    if (!c || !out || n_queries < 0 || (n_queries && !queries)) return COF_E_INVALID;
    if (c->device < 0) return c->fail(COF_E_UNSUPPORTED, "Host-only context cannot search.");
    (void)hipSetDevice(c->device);
    std::vector<std::shared_ptr<Query>> qs;
    for (int i = 0; i < n_queries; i++) {
        if (!queries[i]) return c->fail(COF_E_INVALID, "null query");
        qs.push_back(queries[i]->q);
    }
    return c->search(qs, mode, out);
}

int cof_stats_get(const cof_ctx *c, cof_stats *out) {
    if (!c || !out) return COF_E_INVALID;
    *out = c->lane.stats;
    out->t_upload_ms = c->t_upload_ms;
    return COF_OK;
}

int cof_tuner_info(cof_ctx *c, char **out, size_t *len) {
    if (!c || !out || !len) return COF_E_INVALID;
    std::string text;
    {
        std::lock_guard<std::mutex> lk(c->tuner.mu);
        for (auto &kv : c->tuner.best) text += kv.first + "\t" + std::to_string(kv.second) + "\n";
    }
    *out = static_cast<char *>(malloc(text.size() + 1));
    if (!*out) return COF_E_NOMEM;
    memcpy(*out, text.data(), text.size());
    (*out)[text.size()] = 0;
    *len = text.size();
    return COF_OK;
}

int cof_tuner_load(cof_ctx *c, const char *path) {
    if (!c || !path) return COF_E_INVALID;
    std::string err;
    const int rc = c->tuner.load(path, err);
    return rc ? c->fail(rc, err) : COF_OK;
}

int cof_tuner_save(cof_ctx *c, const char *path) {
    if (!c || !path) return COF_E_INVALID;
    std::string err;
    const int rc = c->tuner.save(path, err);
    return rc ? c->fail(rc, err) : COF_OK;
}

int cof_launch_info(const cof_ctx *c, char **out, size_t *len) {
    if (!c || !out || !len) return COF_E_INVALID;
    std::string text;
    char buf[64];
    for (const LaunchRec &lr : c->lane.launches) {
        snprintf(buf, sizeof buf, "\t%d\t%.6f\t%llu\n", lr.variant, lr.ms, (unsigned long long)lr.candidates);
        text += lr.kernel + "\t" + lr.sig + buf;
    }
    *out = static_cast<char *>(malloc(text.size() + 1));
    if (!*out) return COF_E_NOMEM;
    memcpy(*out, text.data(), text.size());
    (*out)[text.size()] = 0;
    *len = text.size();
    return COF_OK;
}

int64_t cof_result_count(const cof_result *r) { return r ? (int64_t)r->hits.size() : COF_E_INVALID; }

int cof_result_records(const cof_result *r, cof_hit *out, size_t cap) {
    if (!r || (!out && cap)) return COF_E_INVALID;
    // query by query (each query's run is sorted; the runs lie in launch order)
    size_t at = 0;
    for (size_t q = 0; q < r->queries.size() && at < cap; q++) {
        const size_t b = r->qbegin[q], n = std::min(cap - at, r->qend[q] - b);
        parallel_for(n, 1 << 20, [&](uint64_t x, uint64_t y) {
            memcpy(out + at + x, r->hits.data() + b + x, (y - x) * sizeof(cof_hit));
        });
        at += n;
    }
    return COF_OK;
}

static int format_query(cof_ctx *c, const cof_result *r, int qi, std::string &text, SummaryTable *sum) {
    if (!c || !r || qi < 0 || (size_t)qi >= r->queries.size()) return COF_E_INVALID;
    // hits index the genome they were searched on
    if (r->genome_serial != c->genome_serial || c->genome.files.empty())
        return c->fail(COF_E_INVALID, "The result was searched on another genome than the context's.");
    const size_t b = r->qbegin[qi], e = r->qend[qi];
    static_assert(sizeof(cof_hit) == sizeof(DevHit), "hit layout");
    format_hits(c->genome, *r->queries[qi], reinterpret_cast<const DevHit *>(r->hits.data()) + b, e - b, text, sum);
    return COF_OK;
}

int cof_result_format(cof_ctx *c, const cof_result *r, int qi, const char *path, int append,
                      const char *summary_path) {
    if (!path) return COF_E_INVALID;
    std::string text;
    SummaryTable sum;
    const bool want_sum = summary_path && summary_path[0];
    if (!append) text = header_lines();
    int rc = format_query(c, r, qi, text, want_sum ? &sum : nullptr);
    if (rc) return rc;
    const bool to_stdout = strcmp(path, "-") == 0;
    FILE *f = to_stdout ? stdout : fopen(path, append ? "ab" : "wb");
    if (!f) return c->fail(COF_E_IO, std::string("cannot open ") + path);
    fwrite(text.data(), 1, text.size(), f);
    if (to_stdout) fflush(f);
    else if (fclose(f) != 0) return c->fail(COF_E_IO, std::string("write failed: ") + path);
    if (want_sum) {
        std::string s = summary_text(sum);
        const bool so = strcmp(summary_path, "-") == 0;
        FILE *g = so ? stdout : fopen(summary_path, "wb");
        if (!g) return c->fail(COF_E_IO, std::string("cannot open ") + summary_path);
        fwrite(s.data(), 1, s.size(), g);
        if (so) fflush(g);
        else fclose(g);
    }
    return COF_OK;
}

int cof_result_format_mem(cof_ctx *c, const cof_result *r, int qi, char **out, size_t *len) {
    if (!out || !len) return COF_E_INVALID;
    std::string text;
    int rc = format_query(c, r, qi, text, nullptr);
    if (rc) return rc;
    char *p = (char *)malloc(text.size() + 1);
    if (!p) return COF_E_NOMEM;
    memcpy(p, text.data(), text.size());
    p[text.size()] = 0;
    *out = p;
    *len = text.size();
    return COF_OK;
}

int cof_format_hits(cof_ctx *c, const cof_query *q, const cof_hit *hits, size_t n, char **out, size_t *len) {
    if (!c || !q || !out || !len || (n && !hits)) return COF_E_INVALID;
    const Query &qq = *q->q;
    for (size_t i = 0; i < n; i++)
        if (hits[i].key >= qq.keys.size() || hits[i].pos + qq.P > c->genome.total)
            return c->fail(COF_E_INVALID, "hit out of range");
    std::string text;
    format_hits(c->genome, qq, reinterpret_cast<const DevHit *>(hits), n, text, nullptr);
    char *p = (char *)malloc(text.size() + 1);
    if (!p) return COF_E_NOMEM;
    memcpy(p, text.data(), text.size());
    p[text.size()] = 0;
    *out = p;
    *len = text.size();
    return COF_OK;
}

int cof_result_free(cof_result *r) {
    if (!r) return COF_OK;
    // a large hit array (huge pages to unmap: ~0.25 ms for 16 M decoded
    // hits) is released on a short-lived thread, so that a caller's search
    // loop does not wait for it; a result holds host memory only
    if (r->hits.capacity() * sizeof(cof_hit) >= (size_t(64) << 20)) {
        try {
            std::thread([r] { delete r; }).detach();
            return COF_OK;
        } catch (const std::system_error &) {
            // no thread: released here
        }
    }
    delete r;
    return COF_OK;
}

void cof_free(void *p) { free(p); }

}  // extern "C"
