// Genome loading, packing and synthetic generation (host side).
#include "genome.hpp"

#include <dirent.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <thread>

namespace cof {

// ------------------------------------------------------------------ threads

int n_host_threads() {
    const char *e = getenv("COF_HOST_THREADS");
    if (e && atoi(e) > 0) return atoi(e);
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    return (int)std::min(n, 16u);
}

void parallel_for(uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)> &fn, int max_threads) {
    if (n == 0) return;
    uint64_t nt = (uint64_t)n_host_threads();
    if (max_threads > 0) nt = std::min<uint64_t>(nt, (uint64_t)max_threads);
    if (grain == 0) grain = 1;
    nt = std::min<uint64_t>(nt, (n + grain - 1) / grain);
    if (nt <= 1) { fn(0, n); return; }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (uint64_t t = 0; t < nt; t++) {
        uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto &x : th) x.join();
}

// --------------------------------------------------------------- HostGenome

static const uint8_t *code_table() {
    static uint8_t t[256];
    static bool init = false;
    if (!init) {
        for (int i = 0; i < 256; i++) t[i] = 4;
        t[(unsigned char)'T'] = 0; t[(unsigned char)'C'] = 1;
        t[(unsigned char)'A'] = 2; t[(unsigned char)'G'] = 3;
        init = true;
    }
    return t;
}

void HostGenome::clear() {
    total = 0;
    lo.clear(); hi.clear(); ot.clear();
    exc.clear(); poison.clear(); files.clear();
    cur_lo_ = cur_hi_ = cur_ot_ = 0;
}

void HostGenome::flush_word() {
    lo.push_back(cur_lo_); hi.push_back(cur_hi_); ot.push_back(cur_ot_);
    cur_lo_ = cur_hi_ = cur_ot_ = 0;
}

void HostGenome::add_exc(uint64_t pos, char c) {
    if (!exc.empty()) {
        ExcRun &r = exc.back();
        if (r.ch == c && r.start + r.len == pos) { r.len++; goto poisoned; }
    }
    exc.push_back({pos, 1, c});
poisoned:
    if (c == ';') poison.push_back(pos);
}

void HostGenome::append(const char *s, size_t n) {
    const uint8_t *tab = code_table();
    for (size_t i = 0; i < n; i++) {
        const uint32_t bit = (uint32_t)(total & 31);
        const uint8_t code = tab[(unsigned char)s[i]];
        if (code < 4) {
            cur_lo_ |= (uint32_t)(code & 1) << bit;
            cur_hi_ |= (uint32_t)(code >> 1) << bit;
        } else {
            cur_ot_ |= 1u << bit;
            add_exc(total, s[i]);
        }
        total++;
        if (bit == 31) flush_word();
    }
}

void HostGenome::append_parallel(const char *s, size_t n) {
    const int nt = n_host_threads();
    if (n < (size_t(1) << 22) || nt <= 1) {
        append(s, n);
        return;
    }
    size_t i = 0;
    while (i < n && (total & 31)) append(s + i++, 1);  // up to a word boundary
    const uint64_t nw = (n - i) / 32, w0 = total >> 5;
    lo.resize(w0 + nw);
    hi.resize(w0 + nw);
    ot.resize(w0 + nw);
    const char *body = s + i;
    const uint8_t *tab = code_table();
    struct Part {
        std::vector<ExcRun> exc;
        std::vector<uint64_t> poison;
    };
    std::vector<Part> parts((size_t)nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        th.emplace_back([&, t] {
            Part &pt = parts[(size_t)t];
            const uint64_t wb = nw * (uint64_t)t / (uint64_t)nt, we = nw * (uint64_t)(t + 1) / (uint64_t)nt;
            for (uint64_t w = wb; w < we; w++) {
                uint32_t l = 0, h = 0, o = 0;
                const char *c = body + 32 * w;
                for (uint32_t b = 0; b < 32; b++) {
                    const uint8_t code = tab[(unsigned char)c[b]];
                    if (code < 4) {
                        l |= (uint32_t)(code & 1) << b;
                        h |= (uint32_t)(code >> 1) << b;
                    } else {
                        o |= 1u << b;
                        const uint64_t pos = total + 32 * w + b;
                        if (!pt.exc.empty() && pt.exc.back().ch == c[b] && pt.exc.back().start + pt.exc.back().len == pos)
                            pt.exc.back().len++;
                        else
                            pt.exc.push_back({pos, 1, c[b]});
                        if (c[b] == ';') pt.poison.push_back(pos);
                    }
                }
                lo[w0 + w] = l;
                hi[w0 + w] = h;
                ot[w0 + w] = o;
            }
        });
    }
    for (auto &x : th) x.join();
    for (Part &pt : parts) {  // chunks in position order: runs may continue across chunks
        for (const ExcRun &r : pt.exc) {
            if (!exc.empty() && exc.back().ch == r.ch && exc.back().start + exc.back().len == r.start)
                exc.back().len += r.len;
            else
                exc.push_back(r);
        }
        poison.insert(poison.end(), pt.poison.begin(), pt.poison.end());
    }
    total += 32 * nw;
    append(body + 32 * nw, n - i - 32 * nw);
}

void HostGenome::begin_file(const std::string &path) {
    if (!files.empty()) append_char(';');  // separator between files
    GenomeFile f;
    f.path = path;
    f.base = total;
    files.push_back(f);
}

void HostGenome::end_file() { files.back().len = total - files.back().base; }

void HostGenome::finish() {
    if (total & 31) {
        lo.push_back(cur_lo_); hi.push_back(cur_hi_); ot.push_back(cur_ot_);
        cur_lo_ = cur_hi_ = cur_ot_ = 0;
    }
    lo.resize(nwords() + kPadWords, 0);
    hi.resize(nwords() + kPadWords, 0);
    ot.resize(nwords() + kPadWords, 0);
}

char HostGenome::char_at(uint64_t pos) const {
    const uint64_t w = pos >> 5;
    const uint32_t b = (uint32_t)(pos & 31);
    if ((ot[w] >> b) & 1) {
        auto it = std::upper_bound(exc.begin(), exc.end(), pos,
                                   [](uint64_t p, const ExcRun &r) { return p < r.start; });
        if (it == exc.begin()) return 'N';
        --it;
        return (pos < it->start + it->len) ? it->ch : 'N';
    }
    static const char b2c[4] = {'T', 'C', 'A', 'G'};
    return b2c[((lo[w] >> b) & 1) | (((hi[w] >> b) & 1) << 1)];
}

void HostGenome::text(uint64_t pos, uint64_t n, char *out) const {
    for (uint64_t i = 0; i < n; i++) out[i] = char_at(pos + i);
}

int HostGenome::file_of(uint64_t pos) const {
    auto it = std::upper_bound(files.begin(), files.end(), pos,
                               [](uint64_t p, const GenomeFile &f) { return p < f.base; });
    if (it == files.begin()) return -1;
    --it;
    if (pos >= it->base + it->len) return -1;
    return (int)(it - files.begin());
}

uint64_t HostGenome::n_contigs() const {
    uint64_t n = 0;
    for (auto &f : files) n += f.names.size();
    return n;
}

// ------------------------------------------------------------------ readers

// The file a reader opens for `path`: when `path` is a symbolic link, the
// link's raw target text (read_fasta.cpp:17-22, read_twobit.cpp:58-63), so a
// relative target resolves against the working directory, not the link's
// directory; otherwise `path` itself.
static FILE *open_genome_file(const std::string &path) {
    char buf[PATH_MAX + 1];
    const ssize_t n = readlink(path.c_str(), buf, PATH_MAX);
    if (n >= 0) {
        buf[n] = 0;
        return fopen(buf, "rb");
    }
    return fopen(path.c_str(), "rb");
}

// FASTA as read_fasta.cpp:11-57 holds it: lines split on '\n', one trailing
// '\r' dropped, empty lines skipped, '>' lines start a record (name = rest of
// the line, ';' before every record but the first), other lines uppercased
// (C-locale toupper) and appended.
int load_fasta(const std::string &path, HostGenome &g, std::string &err) {
    FILE *f = open_genome_file(path);
    if (!f) return 1;
    int c0 = fgetc(f);
    if (c0 != '>') { fclose(f); return 1; }
    struct stat stt;
    if (fstat(fileno(f), &stt) != 0) { fclose(f); err = "cannot stat " + path; return -1; }
    const size_t size = (size_t)stt.st_size;
    std::vector<char> file(size);
    rewind(f);
    if (size && fread(file.data(), 1, size, f) != size) { fclose(f); err = "read error: " + path; return -1; }
    fclose(f);
    // chunks that start at line starts; pass 1 counts each chunk's output
    // bytes (sequence bytes and the ';' before every record but the first)
    const int nt = std::max(1, std::min<int>(n_host_threads(), (int)(size >> 20) + 1));
    std::vector<size_t> cb((size_t)nt + 1, size);
    cb[0] = 0;
    for (int t = 1; t < nt; t++) {
        size_t b = size * (size_t)t / (size_t)nt;
        const char *nl = b < size ? static_cast<const char *>(memchr(file.data() + b, '\n', size - b)) : nullptr;
        cb[(size_t)t] = nl ? (size_t)(nl - file.data()) + 1 : size;
    }
    for (int t = 1; t <= nt; t++) cb[(size_t)t] = std::max(cb[(size_t)t], cb[(size_t)t - 1]);
    struct Rec {
        uint64_t out;  // output offset (in the chunk) where the record's sequence starts
        std::string name;
    };
    struct Chunk {
        uint64_t out = 0;
        std::vector<Rec> recs;
    };
    std::vector<Chunk> ch((size_t)nt);
    // for_lines(t, fn): fn(line start, length without '\n' and trailing '\r')
    auto for_lines = [&](int t, auto &&fn) {
        size_t p = cb[(size_t)t];
        const size_t e = cb[(size_t)t + 1];
        while (p < e) {
            const char *nl = static_cast<const char *>(memchr(file.data() + p, '\n', e - p));
            const size_t le = nl ? (size_t)(nl - file.data()) : e;
            size_t n = le - p;
            if (n > 0 && file[le - 1] == '\r') n--;
            if (n > 0) fn(p, n);
            p = le + 1;
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                Chunk &c = ch[(size_t)t];
                for_lines(t, [&](size_t p, size_t n) {
                    if (file[p] == '>') {
                        if (p != 0) c.out++;  // ';' before every record but the first
                        c.recs.push_back({c.out, std::string(file.data() + p + 1, n - 1)});
                    } else {
                        c.out += n;
                    }
                });
            });
        for (auto &x : th) x.join();
    }
    std::vector<uint64_t> base((size_t)nt + 1, 0);
    for (int t = 0; t < nt; t++) base[(size_t)t + 1] = base[(size_t)t] + ch[(size_t)t].out;
    // pass 2: uppercased sequence bytes and separators, at the chunks' offsets
    std::vector<char> text(base[(size_t)nt]);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                char *o = text.data() + base[(size_t)t];
                for_lines(t, [&](size_t p, size_t n) {
                    if (file[p] == '>') {
                        if (p != 0) *o++ = ';';
                    } else {
                        for (size_t k = 0; k < n; k++) {
                            const char c = file[p + k];
                            *o++ = (c >= 'a' && c <= 'z') ? (char)(c - 32) : c;
                        }
                    }
                });
            });
        for (auto &x : th) x.join();
    }
    std::vector<char>().swap(file);
    g.begin_file(path);
    GenomeFile &gf = g.files.back();
    const uint64_t off = g.total - gf.base;  // file text offset of text[0]
    for (int t = 0; t < nt; t++)
        for (Rec &r : ch[(size_t)t].recs) {
            gf.chrpos.push_back(off + base[(size_t)t] + r.out);
            gf.names.push_back(std::move(r.name));
        }
    g.append_parallel(text.data(), text.size());
    g.end_file();
    return 0;
}

static bool rd32(FILE *f, uint32_t &v) { return fread(&v, 4, 1, f) == 1; }

// UCSC .2bit v0 as read_twobit.cpp:44-131 holds it: records read in file
// order after the index, bases T/C/A/G MSB-first, N blocks as 'N', mask blocks
// ignored, ';' between records.
int load_twobit(const std::string &path, HostGenome &g, std::string &err) {
    FILE *f = open_genome_file(path);
    if (!f) return 1;
    uint32_t magic, ver, cnt, tmp;
    if (!rd32(f, magic) || magic != 0x1A412743u) { fclose(f); return 1; }
    if (!rd32(f, ver) || ver != 0) { fclose(f); return 1; }
    if (!rd32(f, cnt) || !rd32(f, tmp)) { fclose(f); err = "truncated 2bit header"; return -1; }
    std::vector<std::string> names(cnt);
    for (uint32_t i = 0; i < cnt; i++) {
        unsigned char nl = 0;
        if (fread(&nl, 1, 1, f) != 1) { fclose(f); err = "truncated 2bit index"; return -1; }
        std::string nm(nl, '\0');
        if (nl && fread(&nm[0], 1, nl, f) != nl) { fclose(f); err = "truncated 2bit index"; return -1; }
        names[i] = nm;
        if (!rd32(f, tmp)) { fclose(f); err = "truncated 2bit index"; return -1; }
    }
    g.begin_file(path);
    GenomeFile *gf = &g.files.back();
    static const char b2c[4] = {'T', 'C', 'A', 'G'};
    std::vector<unsigned char> raw;
    std::vector<char> seq;
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t len = 0, nb = 0, mb = 0;
        rd32(f, len);
        rd32(f, nb);
        std::vector<uint32_t> ns(nb), nz(nb);
        for (uint32_t j = 0; j < nb; j++) rd32(f, ns[j]);
        for (uint32_t j = 0; j < nb; j++) rd32(f, nz[j]);
        rd32(f, mb);
        fseek(f, (long)mb * 8 + 4, SEEK_CUR);
        const uint64_t nraw = (uint64_t)len / 4 + ((len & 3) ? 1 : 0);
        raw.assign(nraw, 0);
        size_t got = nraw ? fread(raw.data(), 1, nraw, f) : 0;
        (void)got;
        seq.resize(len);
        parallel_for(len, 1 << 20, [&](uint64_t b, uint64_t e) {
            for (uint64_t p = b; p < e; p++) seq[p] = b2c[(raw[p >> 2] >> ((3 - (p & 3)) * 2)) & 3];
        });
        for (uint32_t j = 0; j < nb; j++) {
            uint64_t s = ns[j], e = std::min<uint64_t>((uint64_t)ns[j] + nz[j], len);
            for (uint64_t p = s; p < e; p++) seq[p] = 'N';
        }
        gf->chrpos.push_back(g.total - gf->base);
        gf->names.push_back(names[i]);
        g.append_parallel(seq.data(), seq.size());
        if (i + 1 < cnt) g.append_char(';');
    }
    fclose(f);
    g.end_file();
    return 0;
}

static int load_one(const std::string &p, HostGenome &g, int &nfiles, std::string &err, bool verbose) {
    if (verbose) std::cerr << "Reading " << p << "..." << std::endl;
    int r = load_fasta(p, g, err);
    if (r == 1) r = load_twobit(p, g, err);
    if (r < 0) return -1;
    if (r == 1) {
        if (verbose) std::cerr << "Skipping non-acceptable file " << p << "..." << std::endl;
        return 0;
    }
    nfiles++;
    return 0;
}

int load_path(const std::string &path, HostGenome &g, int &nfiles, std::string &err, bool verbose) {
    g.clear();
    nfiles = 0;
    DIR *d = opendir(path.c_str());
    if (!d) {
        struct stat st;
        if (stat(path.c_str(), &st) != 0) {
            err = "An error has occured while opening directory: " + path;
            return COF_E_IO;
        }
        if (load_one(path, g, nfiles, err, verbose)) return COF_E_IO;
    } else {
        struct dirent *ent;
        while ((ent = readdir(d)) != nullptr) {
            if (ent->d_type == DT_REG || ent->d_type == DT_LNK) {
                std::string p = path + "/" + ent->d_name;
                if (load_one(p, g, nfiles, err, verbose)) { closedir(d); return COF_E_IO; }
            }
        }
        closedir(d);
    }
    g.finish();
    return COF_OK;
}

// --------------------------------------------------------------- synthetic

static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() { s += 0x9E3779B97F4A7C15ull; return mix64(s); }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

struct SynthLayout {
    std::vector<uint64_t> len, gstart;                        // per contig
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> nruns;  // merged [s, e) per contig
    uint64_t total = 0;
    uint32_t thr[3];
    std::vector<uint64_t> cseed;
};

static void merge_runs(std::vector<std::pair<uint64_t, uint64_t>> &v) {
    std::sort(v.begin(), v.end());
    std::vector<std::pair<uint64_t, uint64_t>> o;
    for (auto &r : v) {
        if (!o.empty() && r.first <= o.back().second) o.back().second = std::max(o.back().second, r.second);
        else o.push_back(r);
    }
    v.swap(o);
}

// Pinned generator:
//  contig lengths: mode 0 weights 0.5+U[0,1), mode 1 exp(U[log min, log max]);
//    floor(total*w/sum) for all but the last, which takes the remainder;
//  base at offset p of contig c: 16-bit field (p&3) of mix64(cseed_c + p/4),
//    cut at T < (1-gc)/2, C < +gc/2, A < +(1-gc)/2, else G (x 65536, rounded);
//  N runs: round(n_frac*len/mean) runs, length U{min..max}, start U{0..len-rl},
//    overlaps merged;  lowercase runs (FASTA/2bit mask only): alternating
//    unmasked / masked runs with means 5000*(1-f)/f and 5000.
static SynthLayout synth_layout(const cof_synth_spec &sp) {
    SynthLayout L;
    const uint32_t n = sp.n_contigs ? sp.n_contigs : 1;
    Rng rng(mix64(sp.seed ^ 0xC0F5EEDull));
    std::vector<double> w(n);
    double sum = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (sp.length_mode == 1) {
            double a = std::log((double)std::max<uint64_t>(sp.min_len, 1));
            double b = std::log((double)std::max<uint64_t>(sp.max_len, sp.min_len ? sp.min_len : 1));
            w[i] = std::exp(a + rng.uni() * (b - a));
        } else {
            w[i] = 0.5 + rng.uni();
        }
        sum += w[i];
    }
    L.len.resize(n);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (i + 1 < n) L.len[i] = (uint64_t)std::floor((double)sp.total_len * w[i] / sum);
        else L.len[i] = sp.total_len - acc;
        acc += L.len[i];
    }
    L.gstart.resize(n);
    uint64_t g = 0;
    for (uint32_t i = 0; i < n; i++) {
        L.gstart[i] = g;
        g += L.len[i] + (i + 1 < n ? 1 : 0);
    }
    L.total = g;
    const double gc = sp.gc;
    L.thr[0] = (uint32_t)std::llround(65536.0 * (1 - gc) / 2);
    L.thr[1] = L.thr[0] + (uint32_t)std::llround(65536.0 * gc / 2);
    L.thr[2] = L.thr[1] + (uint32_t)std::llround(65536.0 * (1 - gc) / 2);
    L.cseed.resize(n);
    L.nruns.resize(n);
    for (uint32_t c = 0; c < n; c++) {
        L.cseed[c] = mix64(sp.seed * 0x100000001B3ull + c + 1);
        Rng r(mix64(L.cseed[c] ^ 0xABCDEFull));
        const uint64_t len = L.len[c];
        const double mean = 0.5 * (sp.n_run_min + sp.n_run_max);
        if (sp.n_frac > 0 && mean > 0 && len > 0) {
            uint64_t nr = (uint64_t)std::llround(sp.n_frac * (double)len / mean);
            for (uint64_t k = 0; k < nr; k++) {
                uint64_t rl = sp.n_run_min + r.below((uint64_t)sp.n_run_max - sp.n_run_min + 1);
                if (rl > len) rl = len;
                uint64_t st = r.below(len - rl + 1);
                L.nruns[c].push_back({st, st + rl});
            }
            merge_runs(L.nruns[c]);
        }
    }
    return L;
}

static inline uint32_t synth_code(const SynthLayout &L, uint32_t c, uint64_t p) {
    const uint64_t h = mix64(L.cseed[c] + (p >> 2));
    const uint32_t u = (uint32_t)(h >> (16 * (p & 3))) & 0xFFFF;
    return u < L.thr[0] ? 0 : u < L.thr[1] ? 1 : u < L.thr[2] ? 2 : 3;
}

void synth_genome(const cof_synth_spec &sp, HostGenome &g) {
    g.clear();
    SynthLayout L = synth_layout(sp);
    const uint32_t n = (uint32_t)L.len.size();
    g.files.push_back(GenomeFile());
    GenomeFile &f = g.files.back();
    f.path = "synthetic";
    f.base = 0;
    f.len = L.total;
    for (uint32_t c = 0; c < n; c++) {
        f.names.push_back("chr" + std::to_string(c + 1));
        f.chrpos.push_back(L.gstart[c]);
    }
    g.total = L.total;
    const uint64_t nw = g.nwords();
    g.lo.assign(nw + kPadWords, 0);
    g.hi.assign(nw + kPadWords, 0);
    g.ot.assign(nw + kPadWords, 0);
    // exceptions and poison (sequential: few runs)
    for (uint32_t c = 0; c < n; c++) {
        for (auto &r : L.nruns[c]) g.exc.push_back({L.gstart[c] + r.first, r.second - r.first, 'N'});
        if (c + 1 < n) {
            const uint64_t sep = L.gstart[c] + L.len[c];
            g.exc.push_back({sep, 1, ';'});
            g.poison.push_back(sep);
        }
    }
    // planes, in parallel over words
    parallel_for(nw, 1 << 14, [&](uint64_t wb, uint64_t we) {
        uint64_t p = wb * 32;
        uint32_t c = (uint32_t)(std::upper_bound(L.gstart.begin(), L.gstart.end(), p) - L.gstart.begin()) - 1;
        size_t ri = 0;
        auto seek_run = [&](uint32_t cc, uint64_t off) {
            auto &v = L.nruns[cc];
            ri = (size_t)(std::partition_point(v.begin(), v.end(),
                                               [off](const std::pair<uint64_t, uint64_t> &r) { return r.second <= off; }) -
                          v.begin());
        };
        seek_run(c, p - L.gstart[c]);
        for (uint64_t w = wb; w < we; w++) {
            uint32_t wl = 0, wh = 0, wo = 0;
            for (uint32_t b = 0; b < 32; b++, p++) {
                if (p >= L.total) break;
                while (c + 1 < n && p >= L.gstart[c + 1]) { c++; ri = 0; }
                const uint64_t off = p - L.gstart[c];
                if (off >= L.len[c]) { wo |= 1u << b; continue; }  // separator
                auto &v = L.nruns[c];
                while (ri < v.size() && v[ri].second <= off) ri++;
                if (ri < v.size() && v[ri].first <= off) { wo |= 1u << b; continue; }
                const uint32_t code = synth_code(L, c, off);
                wl |= (code & 1) << b;
                wh |= (code >> 1) << b;
            }
            g.lo[w] = wl; g.hi[w] = wh; g.ot[w] = wo;
        }
    });
}

static std::vector<std::pair<uint64_t, uint64_t>> lower_runs(const cof_synth_spec &sp, uint64_t cseed, uint64_t len) {
    std::vector<std::pair<uint64_t, uint64_t>> v;
    if (sp.lower_frac <= 0) return v;
    const double f = std::min(sp.lower_frac, 0.99);
    Rng r(mix64(cseed ^ 0x5151ull));
    const uint64_t mu = (uint64_t)std::max(1.0, 5000.0 * (1 - f) / f), mm = 5000;
    uint64_t p = 0;
    bool masked = false;
    while (p < len) {
        uint64_t rl = 1 + r.below(2 * (masked ? mm : mu));
        uint64_t e = std::min(len, p + rl);
        if (masked) v.push_back({p, e});
        p = e;
        masked = !masked;
    }
    return v;
}

int synth_write(const cof_synth_spec &sp, const std::string &path, int format, std::string &err) {
    SynthLayout L = synth_layout(sp);
    const uint32_t n = (uint32_t)L.len.size();
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { err = "cannot write " + path; return COF_E_IO; }
    static const char b2c[4] = {'T', 'C', 'A', 'G'};
    if (format == 0) {
        const uint32_t W = sp.line_width ? sp.line_width : 60;
        std::vector<char> seq;
        for (uint32_t c = 0; c < n; c++) {
            fprintf(f, ">chr%u\n", c + 1);
            const uint64_t len = L.len[c];
            seq.resize(len);
            parallel_for(len, 1 << 20, [&](uint64_t b, uint64_t e) {
                for (uint64_t p = b; p < e; p++) seq[p] = b2c[synth_code(L, c, p)];
            });
            for (auto &r : L.nruns[c]) memset(seq.data() + r.first, 'N', r.second - r.first);
            for (auto &r : lower_runs(sp, L.cseed[c], len))
                for (uint64_t p = r.first; p < r.second; p++) seq[p] = (char)(seq[p] + 32);
            std::vector<char> out;
            out.reserve(len + len / W + 2);
            for (uint64_t p = 0; p < len; p += W) {
                uint64_t e = std::min(len, p + W);
                out.insert(out.end(), seq.begin() + p, seq.begin() + e);
                out.push_back('\n');
            }
            fwrite(out.data(), 1, out.size(), f);
        }
    } else {
        // UCSC 2bit v0: header, index, then records in index order.
        std::vector<std::string> names;
        for (uint32_t c = 0; c < n; c++) names.push_back("chr" + std::to_string(c + 1));
        uint64_t off = 16;
        for (auto &nm : names) off += 1 + nm.size() + 4;
        uint32_t hdr[4] = {0x1A412743u, 0, n, 0};
        fwrite(hdr, 4, 4, f);
        std::vector<std::vector<std::pair<uint64_t, uint64_t>>> masks(n);
        for (uint32_t c = 0; c < n; c++) masks[c] = lower_runs(sp, L.cseed[c], L.len[c]);
        for (uint32_t c = 0; c < n; c++) {
            unsigned char nl = (unsigned char)names[c].size();
            fwrite(&nl, 1, 1, f);
            fwrite(names[c].data(), 1, nl, f);
            uint32_t o32 = (uint32_t)off;
            fwrite(&o32, 4, 1, f);
            off += 4 + 4 + 8 * L.nruns[c].size() + 4 + 8 * masks[c].size() + 4 + (L.len[c] + 3) / 4;
        }
        for (uint32_t c = 0; c < n; c++) {
            const uint64_t len = L.len[c];
            uint32_t v = (uint32_t)len;
            fwrite(&v, 4, 1, f);
            v = (uint32_t)L.nruns[c].size();
            fwrite(&v, 4, 1, f);
            for (auto &r : L.nruns[c]) { v = (uint32_t)r.first; fwrite(&v, 4, 1, f); }
            for (auto &r : L.nruns[c]) { v = (uint32_t)(r.second - r.first); fwrite(&v, 4, 1, f); }
            v = (uint32_t)masks[c].size();
            fwrite(&v, 4, 1, f);
            for (auto &r : masks[c]) { v = (uint32_t)r.first; fwrite(&v, 4, 1, f); }
            for (auto &r : masks[c]) { v = (uint32_t)(r.second - r.first); fwrite(&v, 4, 1, f); }
            v = 0;
            fwrite(&v, 4, 1, f);
            std::vector<unsigned char> raw((len + 3) / 4, 0);
            parallel_for(raw.size(), 1 << 18, [&](uint64_t b, uint64_t e) {
                for (uint64_t i = b; i < e; i++) {
                    unsigned char x = 0;
                    for (int k = 0; k < 4; k++) {
                        uint64_t p = i * 4 + k;
                        uint32_t code = p < len ? synth_code(L, c, p) : 0;
                        x |= (unsigned char)(code << ((3 - k) * 2));
                    }
                    raw[i] = x;
                }
            });
            for (auto &r : L.nruns[c])  // N positions store T (0), as faToTwoBit does
                for (uint64_t p = r.first; p < r.second; p++) raw[p >> 2] &= (unsigned char)~(3u << ((3 - (p & 3)) * 2));
            fwrite(raw.data(), 1, raw.size(), f);
        }
    }
    if (fclose(f) != 0) { err = "write failed: " + path; return COF_E_IO; }
    return COF_OK;
}

}  // namespace cof
