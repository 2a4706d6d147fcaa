// Input parsing and compare-key expansion with the reference's semantics
// (cas-offinder.cpp:570-694), and the device encodings of pattern and keys.
#include "query.hpp"

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>

namespace cof {

uint32_t mismatch_classes(char p) {
    constexpr uint32_t T = 1u << kClsT, C = 1u << kClsC, A = 1u << kClsA, G = 1u << kClsG, O = 1u << kClsO;
    switch (p) {
        case 'A': return T | C | G | O;
        case 'C': return T | A | G | O;
        case 'G': return T | C | A | O;
        case 'T': return C | A | G | O;
        case 'R': return C | T;  // A/G
        case 'Y': return A | G;  // C/T
        case 'S': return A | T;  // G/C
        case 'W': return C | G;  // A/T
        case 'K': return A | C;  // G/T
        case 'M': return G | T;  // A/C
        case 'B': return A;      // C/G/T
        case 'D': return C;      // A/G/T
        case 'H': return G;      // A/C/T
        case 'V': return T;      // A/C/G
        default: return 0;       // N and every other byte
    }
}

uint32_t class_of(char c) {
    switch (c) {
        case 'T': return 1u << kClsT;
        case 'C': return 1u << kClsC;
        case 'A': return 1u << kClsA;
        case 'G': return 1u << kClsG;
        default: return 1u << kClsO;
    }
}

static char complement(char c) {
    switch (c) {
        case 'A': return 'T'; case 'T': return 'A';
        case 'G': return 'C'; case 'C': return 'G';
        case 'R': return 'Y'; case 'Y': return 'R';
        case 'M': return 'K'; case 'K': return 'M';
        case 'H': return 'D'; case 'D': return 'H';
        case 'B': return 'V'; case 'V': return 'B';
        default: return c;
    }
}

void revcomp(std::string &s) {
    for (auto &c : s) c = complement(c);
    std::reverse(s.begin(), s.end());
}

namespace {

// std::getline over a byte buffer with the istream flag behaviour the
// reference's parseLine (cas-offinder.cpp:570-582) depends on.
class LineReader {
public:
    LineReader(const char *s, size_t n) : s_(s), n_(n) {}
    bool eof() const { return eof_; }
    bool getline(std::string &out) {
        if (eof_ || fail_) { fail_ = true; return false; }
        out.clear();
        if (pos_ >= n_) { eof_ = fail_ = true; return false; }
        const char *nl = static_cast<const char *>(memchr(s_ + pos_, '\n', n_ - pos_));
        size_t e = nl ? (size_t)(nl - s_) : n_;
        out.assign(s_ + pos_, e - pos_);
        if (nl) pos_ = e + 1; else { pos_ = n_; eof_ = true; }
        return true;
    }

private:
    const char *s_;
    size_t n_, pos_ = 0;
    bool eof_ = false, fail_ = false;
};

struct ParseError {
    std::string msg;
};

// parseLine: 1 read, 0 nothing read; throws on the expected-line EOF case.
bool parse_line(LineReader &in, std::string &line, bool expect_line) {
    if (expect_line && in.eof()) throw ParseError{"Unexpected end of file."};
    if (!in.getline(line)) return false;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
}

// whitespace tokens as istream_iterator<string> yields them (cas-offinder.cpp:20-24)
std::vector<std::string> tokens(const std::string &s) {
    std::vector<std::string> v;
    size_t i = 0;
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; };
    while (i < s.size()) {
        while (i < s.size() && ws(s[i])) i++;
        size_t b = i;
        while (i < s.size() && !ws(s[i])) i++;
        if (i > b) v.emplace_back(s, b, i - b);
    }
    return v;
}

void upper(std::string &s) {
    for (auto &c : s)
        if (c >= 'a' && c <= 'z') c = (char)(c - 32);
}

struct KeyEntry {
    std::string id;
    uint16_t threshold;
    std::vector<BulgeInfo> infos;
};

// add_compare (cas-offinder.cpp:14-18): first (id, threshold) wins, infos accumulate.
void add_key(std::map<std::string, KeyEntry> &m, const std::string &key, const std::string &id, uint16_t t,
             BulgeInfo bi) {
    auto it = m.find(key);
    if (it == m.end()) m.emplace(key, KeyEntry{id, t, {std::move(bi)}});
    else it->second.infos.push_back(std::move(bi));
}

std::string substr_checked(const std::string &s, uint64_t pos, uint64_t n = ~0ull) {
    if (pos > s.size())
        throw ParseError{"basic_string::substr: __pos (which is " + std::to_string(pos) +
                         ") > this->size() (which is " + std::to_string(s.size()) + ")"};
    return s.substr((size_t)pos, (size_t)std::min<uint64_t>(n, s.size() - pos));
}

// Every compare key a guide expands to, in the reference's order
// (cas-offinder.cpp:647-682), with the bulge family it belongs to:
// emit(key, info, shift, split) with shift 0 for the D = 0 key, i for a DNA
// bulge of size i and D + i for an RNA bulge of size i, and split = j.
template <typename Emit>
void expand_guide(const std::string &g, uint32_t D, uint32_t R, bool rev, Emit &&emit) {
    // first / last non-N index as std::string::find_*_not_of returns them
    const uint64_t f = g.find_first_not_of('N'), l = g.find_last_not_of('N');
    if (D == 0) {
        emit(g, BulgeInfo{"0", 0}, 0u, 0u);
    } else {
        const std::string base = rev ? g + std::string(D, 'N') : std::string(D, 'N') + g;
        for (uint32_t i = 1; i <= D; i++) {
            const uint64_t lim = l + 2;  // wraps to 1 for an all-N guide
            for (uint32_t j = (uint32_t)f; (uint64_t)j < lim; j++) {
                std::string key = rev ? g.substr(0, j) + std::string(i, 'N') + substr_checked(g, j) +
                                            std::string(D - i, 'N')
                                      : std::string(D - i, 'N') + g.substr(0, j) + std::string(i, 'N') +
                                            substr_checked(g, j);
                const int idx = rev ? -(int)j : (int)j;
                const bool is_base = key == base;
                emit(key, BulgeInfo{is_base ? "0" : std::to_string(i), idx}, i, j);
            }
        }
    }
    for (uint32_t i = 1; i <= R; i++) {
        const uint64_t lim = l + 1 - i;  // size_t arithmetic, may wrap
        for (uint32_t j = (uint32_t)f; (uint64_t)j < lim; j++) {
            std::string bulge = substr_checked(g, j, i);
            std::string rest = substr_checked(g, (uint64_t)j + i);
            std::string key = rev ? g.substr(0, j) + rest + std::string(D + i, 'N')
                                  : std::string(D + i, 'N') + g.substr(0, j) + rest;
            emit(key, BulgeInfo{bulge, rev ? -(int)j : (int)j}, D + i, j);
        }
    }
}

}  // namespace

int parse_query(const char *text, size_t len, Query &q, std::string &err) {
    q = Query();
    LineReader in(text, len);
    std::map<std::string, KeyEntry> m;
    try {
        std::string line;
        parse_line(in, q.chrdir, true);
        parse_line(in, line, true);
        auto tk = tokens(line);
        if (tk.size() != 1 && tk.size() != 3) throw ParseError{"Malformed input file."};
        if (tk.size() == 3) {
            int d = atoi(tk[1].c_str()), r = atoi(tk[2].c_str());
            if (d < 0 || r < 0) throw ParseError{"Negative bulge size."};
            q.D = (uint32_t)d;
            q.R = (uint32_t)r;
        }
        const bool rev = tk[0][0] != 'N';  // tested before uppercasing (cas-offinder.cpp:619)
        q.rev = rev;
        q.pattern = rev ? tk[0] + std::string(q.D, 'N') : std::string(q.D, 'N') + tk[0];
        upper(q.pattern);
        q.P = (uint32_t)q.pattern.size();

        const uint32_t D = q.D, R = q.R;
        size_t linecnt = 0, entrycnt = 0;
        while (parse_line(in, line, false)) {
            if (line.empty()) break;
            auto t = tokens(line);
            if (t.size() != 2 && t.size() != 3) throw ParseError{"Malformed input file."};
            std::string g = t[0];
            if (g.size() + D != q.pattern.size())
                throw ParseError{"The length of target sequences should match with the length of pattern sequence."};
            upper(g);
            const uint16_t thr = (uint16_t)atoi(t[1].c_str());
            const std::string id = t.size() == 3 ? t[2] : std::to_string(linecnt);
            expand_guide(g, D, R, rev, [&](const std::string &key, BulgeInfo &&bi, uint32_t, uint32_t) {
                add_key(m, key, id, thr, std::move(bi));
            });
            q.guide_lines.push_back(GuideLine{g, id, thr});
            if (entrycnt == 0) entrycnt = t.size();
            else if (entrycnt != t.size())
                throw ParseError{"The number of entries below 2nd line should be consistent."};
            linecnt++;
        }
        q.guides = linecnt;
    } catch (const ParseError &e) {
        err = e.msg;
        return -1;
    }
    q.keys.reserve(m.size());
    for (auto &kv : m) q.keys.push_back(CompareKey{kv.first, kv.second.id, kv.second.threshold, std::move(kv.second.infos)});
    return 0;
}

// Records of `keys` at offset cbase / width nw (bit b of word w <-> key
// position cbase + 32 w + b) in the table format of devtypes.hpp.  With pad8
// every run (care class, or the general record list) is padded to a multiple
// of 8 records with never-hit records (bias 2^30).
void encode_records(const std::vector<KeyRef> &keys, uint32_t P, int cbase, int nw, bool gen, bool pad8,
                    std::vector<uint32_t> &out, uint32_t &units) {
    const int stride = key_stride(nw, gen), bi = key_bias_index(nw, gen);
    std::vector<std::vector<uint32_t>> rec(keys.size(), std::vector<uint32_t>(stride, 0));
    std::vector<std::vector<uint32_t>> care(keys.size(), std::vector<uint32_t>(nw, 0));
    for (size_t n = 0; n < keys.size(); n++) {
        uint32_t *w = rec[n].data();
        const std::string &s = keys[n].key->seq;
        for (uint32_t i = 0; i < P; i++) {
            const uint32_t mm = mismatch_classes(s[i]);
            if (!mm) continue;
            const int b = (int)i - cbase;
            const uint32_t bit = 1u << (b & 31);
            care[n][b >> 5] |= bit;
            if (gen) {
                for (int c = 0; c < 5; c++)
                    if (mm & (1u << c)) w[c * nw + (b >> 5)] |= bit;
            } else {
                // 2-bit code of the key base: T=0, C=1, A=2, G=3
                const uint32_t code = s[i] == 'T' ? 0 : s[i] == 'C' ? 1 : s[i] == 'A' ? 2 : 3;
                if (code & 1) w[b >> 5] |= bit;
                if (code & 2) w[nw + (b >> 5)] |= bit;
            }
        }
        w[bi] = (uint32_t)(-(int64_t)keys[n].key->threshold - 1);
        w[bi + 1] = keys[n].index;
    }
    std::vector<uint32_t> padrec(stride, 0);
    padrec[bi] = 1u << 30;
    auto pad = [&](size_t n) {
        if (pad8)
            for (size_t i = n; i % 8; i++) out.insert(out.end(), padrec.begin(), padrec.end());
    };
    if (gen) {
        for (auto &r : rec) out.insert(out.end(), r.begin(), r.end());
        pad(rec.size());
        units = (uint32_t)(pad8 ? (rec.size() + 7) / 8 * 8 : rec.size());
        return;
    }
    // care classes in order of first appearance
    std::map<std::vector<uint32_t>, std::vector<size_t>> cls;
    std::vector<std::vector<uint32_t>> order;
    for (size_t n = 0; n < keys.size(); n++) {
        auto it = cls.find(care[n]);
        if (it == cls.end()) { order.push_back(care[n]); cls[care[n]].push_back(n); }
        else it->second.push_back(n);
    }
    for (auto &cm : order) {
        const auto &members = cls[cm];
        uint32_t hdr[kClassHdr] = {0, 0, 0, 0};
        for (int w = 0; w < nw; w++) hdr[w] = cm[w];
        hdr[nw] = (uint32_t)(pad8 ? (members.size() + 7) / 8 * 8 : members.size());
        out.insert(out.end(), hdr, hdr + kClassHdr);
        for (size_t n : members) out.insert(out.end(), rec[n].begin(), rec[n].end());
        pad(members.size());
    }
    units = (uint32_t)order.size();
}

int build_key_table(const std::vector<KeyRef> &keys, uint32_t P, KeyTable &t, std::string &err) {
    t = KeyTable();
    int lo = -1, hi = -1;
    for (auto &k : keys)
        for (uint32_t i = 0; i < P; i++) {
            const char c = k.key->seq[i];
            if (!mismatch_classes(c)) continue;
            t.care_sum++;
            if (lo < 0 || (int)i < lo) lo = (int)i;
            if ((int)i > hi) hi = (int)i;
            if (c != 'A' && c != 'C' && c != 'G' && c != 'T') t.gen = true;
        }
    if (lo < 0) lo = hi = 0;
    if (hi - lo + 1 > 2 * 32) { err = "Key span too wide."; return -1; }
    t.cbase = lo;
    t.nw = hi - lo + 1 <= 32 ? 1 : 2;
    t.nkeys = keys.size();
    for (auto &k : keys) t.tmax = std::max<int>(t.tmax, k.key->threshold);
    encode_records(keys, P, t.cbase, t.nw, t.gen, true, t.words, t.units);
    return 0;
}

namespace {
// 2-bit codes (lo bits then hi bits, T=0 C=1 A=2 G=3) of key bases
// k[off, off + span) as a window must hold them to match exactly: a base
// that is not A/C/G/T (N: not comparable) takes every value.  Writes them
// to out and returns their number, or -1 if there are more than cap.
int range_codes(const std::string &k, size_t off, int span, uint32_t *out, int cap) {
    int n = 1;
    out[0] = 0;
    for (int i = 0; i < span; i++) {
        const char c = k[off + (size_t)i];
        const int b = c == 'T' ? 0 : c == 'C' ? 1 : c == 'A' ? 2 : c == 'G' ? 3 : -1;
        const uint32_t lo = 1u << i, hi = 1u << (span + i);
        if (b >= 0) {
            const uint32_t add = (b & 1 ? lo : 0) | (b & 2 ? hi : 0);
            for (int j = 0; j < n; j++) out[j] |= add;
            continue;
        }
        if (4 * n > cap) return -1;
        for (int j = 0; j < n; j++) {
            out[n + j] = out[j] | lo;
            out[2 * n + j] = out[j] | hi;
            out[3 * n + j] = out[j] | lo | hi;
        }
        n *= 4;
    }
    return n;
}
}  // namespace

bool build_seed_index(const std::vector<KeyRef> &keys, uint32_t P, const KeyTable &t, SeedIndex &s) {
    (void)P;
    s = SeedIndex();
    if (t.gen || t.nw != 1 || t.units != 1 || keys.empty() || keys.size() > 4096) return false;
    const uint32_t care = t.words[0];
    const int C = __builtin_popcount(care);
    const int parts = t.tmax + 1;
    if (parts > kSeedMaxParts || parts > C) return false;
    // ranges: consecutive groups of comparable positions, as equal as possible
    std::vector<int> pos;
    for (int b = 0; b < 32; b++)
        if ((care >> b) & 1) pos.push_back(b);
    SeedDesc &d = s.desc;
    d.parts = (uint32_t)parts;
    d.nrec = (uint32_t)keys.size();
    uint32_t obase = 0;
    for (int p = 0; p < parts; p++) {
        const int g0 = (int)((int64_t)C * p / parts), g1 = (int)((int64_t)C * (p + 1) / parts);
        const int a = pos[g0], z = g1 < C ? pos[g1] : pos[C - 1] + 1;  // bits [a, z)
        const int len = z - a;
        if (len > 5) return false;  // at most 1024 codes per range
        d.first[p] = (uint32_t)a | (uint32_t)len << 8;
        d.mask[p] = ((1u << len) - 1) << a;
        d.obase[p] = obase;
        obase += (1u << (2 * len)) + 1;
    }
    d.noffs = obase;
    // ids per (range, code) by counting sort over the keys' codes (keys
    // expanded over their N positions), in key order within a code
    std::vector<uint32_t> offs(obase + 1, 0);
    std::vector<uint16_t> ids;
    uint32_t cs[64];
    for (int pass = 0; pass < 2; pass++) {
        for (size_t n = 0; n < keys.size(); n++) {
            const std::string &k = keys[n].key->seq;
            for (int p = 0; p < parts; p++) {
                const int a = (int)(d.first[p] & 0xFF), len = (int)(d.first[p] >> 8);
                const int nc = range_codes(k, (size_t)(t.cbase + a), len, cs, 64);
                if (nc < 0) return false;
                for (int j = 0; j < nc; j++) {
                    if (pass == 0) offs[d.obase[p] + cs[j] + 1]++;
                    else ids[offs[d.obase[p] + cs[j]]++] = (uint16_t)n;
                }
            }
        }
        if (pass == 0) {
            for (uint32_t i = 0; i < obase; i++) offs[i + 1] += offs[i];
            ids.resize(offs[obase]);
        } else {  // offs[i] advanced to the end of list i: shift back to starts
            for (uint32_t i = obase; i > 0; i--) offs[i] = offs[i - 1];
            offs[0] = 0;
        }
    }
    offs.resize(obase);
    if (ids.size() > 0xFFFF) return false;
    d.nids = (uint32_t)ids.size();
    std::vector<uint16_t> all;
    for (uint32_t o : offs) all.push_back((uint16_t)o);
    all.insert(all.end(), ids.begin(), ids.end());
    if (all.size() & 1) all.push_back(0);
    s.words.resize(all.size() / 2);
    memcpy(s.words.data(), all.data(), all.size() * 2);
    s.lds_bytes = d.nrec * 12 + (uint32_t)all.size() * 2;
    return s.lds_bytes <= 24 * 1024;
}

bool build_mask_index(const std::vector<KeyRef> &keys, const KeyTable &t, int minex, MaskIndex &s) {
    s = MaskIndex();
    const size_t n = keys.size();
    if (t.gen || t.nw != 1 || t.units != 1 || n == 0 || n > 32u * kMaskMaxWords || (minex != 1 && minex != 2))
        return false;
    const uint32_t care = t.words[0];
    std::vector<int> pos;
    for (int b = 0; b < 32; b++)
        if ((care >> b) & 1) pos.push_back(b);
    const int C = (int)pos.size();
    MaskDesc &d = s.desc;
    // ranges: consecutive groups of comparable positions, as equal as
    // possible, each spanning at most kMaskMaxSpan bits (holes included)
    auto fits = [&](int parts) {
        if (parts > C) return false;
        for (int p = 0; p < parts; p++) {
            const int g0 = (int)((int64_t)C * p / parts), g1 = (int)((int64_t)C * (p + 1) / parts);
            if (g1 <= g0 || pos[g1 - 1] + 1 - pos[g0] > kMaskMaxSpan) return false;
        }
        return true;
    };
    int parts = t.tmax + minex;
    while (parts <= kSeedMaxParts && !fits(parts)) parts++;
    if (parts > kSeedMaxParts) return false;
    d.minex = (uint32_t)minex;
    d.nrec = (uint32_t)n;
    // records in key order, or, when minex 2 meets differing thresholds, by
    // threshold (so that words hold keys of one threshold and each word needs
    // parts - its largest threshold exact ranges: minex 3)
    std::vector<size_t> order(n);
    for (size_t r = 0; r < n; r++) order[r] = r;
    bool mixed = false;
    for (size_t r = 1; r < n; r++) mixed = mixed || keys[r].key->threshold != keys[0].key->threshold;
    if (minex == 2 && mixed) {
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return keys[x].key->threshold > keys[y].key->threshold; });
        d.minex = 3;
    }
    d.care = care;
    d.nk = n <= 32 ? 1 : n <= 64 ? 2 : (uint32_t)(n + 127) / 128 * 4;  // 1, 2 or a multiple of 4
    if (minex == 1 && d.nk > 8) return false;
    if (d.minex == 3 && d.nk >= (uint32_t)kMaskWideWords) {
        // per-word counts of wide tables may use up to kMaskMaxParts ranges: more, shorter
        // ranges raise each word's need (parts - its threshold) and cut the
        // records that reach the exact count, at the price of more row reads
        // and a wider count per word.  Pick the even part count (the kernel
        // reads and counts an even number of rows) of least estimated cost
        // per candidate: rows read (LDS issue of a 16-byte chunk ~8 cycles),
        // the per-word counts, and the expected survivors of a random window
        // (a range of c positions matches with probability 4^-c), each
        // ~60 cycles through the divergent exact-count loop.
        std::vector<int> tw(d.nk, -1);
        for (size_t r = 0; r < n; r++) tw[r / 32] = std::max<int>(tw[r / 32], keys[order[r]].key->threshold);
        auto cost = [&](int P) {
            double pm[kMaskMaxParts + 1] = {1.0};  // distribution of exactly matched ranges
            for (int p = 0; p < P; p++) {
                const int c = (int)((int64_t)C * (p + 1) / P) - (int)((int64_t)C * p / P);
                const double q = 1.0 / (double)(1u << (2 * c));
                for (int k = p + 1; k >= 0; k--) pm[k] = pm[k] * (1 - q) + (k ? pm[k - 1] * q : 0.0);
            }
            double surv = 0;
            for (size_t r = 0; r < n; r++) {
                const int need = P - tw[r / 32];
                for (int k = std::max(need, 0); k <= P; k++) surv += pm[k];
            }
            const int np = P <= 4 ? 4 : (P + 1) / 2 * 2;
            const double count_ops = np <= 4 ? 6 : np <= 6 ? 11 : np <= 8 ? 15 : np <= 10 ? 18 : 22;
            return d.nk * count_ops + 8.0 * np * ((d.nk + 3) / 4) + 60.0 * surv;
        };
        int best = parts;
        const char *force = getenv("COF_MASK_PARTS");  // tests: a given part count where it fits
        if (force && atoi(force) >= parts && atoi(force) <= kMaskMaxParts && fits(atoi(force))) {
            best = atoi(force);
        } else {
            for (int P = parts + 1; P <= kMaskMaxParts; P++)
                if (P % 2 == 0 && fits(P) && cost(P) < cost(best)) best = P;
        }
        parts = best;
    }
    d.parts = (uint32_t)parts;
    uint32_t rows = 0, rbase[kMaskMaxParts];
    for (int p = 0; p < parts; p++) {
        const int g0 = (int)((int64_t)C * p / parts), g1 = (int)((int64_t)C * (p + 1) / parts);
        const int a = pos[g0], span = pos[g1 - 1] + 1 - a;
        rbase[p] = rows;
        d.first[p] = (uint32_t)a | (uint32_t)span << 8 | rows << 16;
        rows += 1u << (2 * span);
    }
    d.zrow = rows++;
    for (int p = parts; p < kMaskMaxParts; p++) d.first[p] = d.zrow << 16;  // span 0: the zero row
    d.recw = (rows * d.nk + 3) & ~3u;
    d.nwords = d.recw + 4 * d.nrec;
    s.words.assign(d.nwords, 0);
    for (uint32_t w = 0; w < d.nk; w++) {
        int tw = 0;
        for (size_t r = 32 * (size_t)w; r < std::min<size_t>(n, 32 * (size_t)w + 32); r++)
            tw = std::max<int>(tw, keys[order[r]].key->threshold);
        const uint32_t need = 32 * (size_t)w < n ? (uint32_t)(parts - tw) : 15u;  // an empty word keeps nothing
        d.need[w / 8] |= need << (4 * (w % 8));
    }
    for (size_t r = 0; r < n; r++) {
        const size_t kr = order[r];  // the key in record slot r
        const std::string &k = keys[kr].key->seq;
        for (int p = 0; p < parts; p++) {
            const int a = (int)(d.first[p] & 0xFF), span = (int)((d.first[p] >> 8) & 0xFF);
            uint32_t cs[1u << (2 * kMaskMaxSpan)];
            const int nc = range_codes(k, (size_t)(t.cbase + a), span, cs, 1 << (2 * kMaskMaxSpan));
            for (int j = 0; j < nc; j++)
                s.words[mask_word(d.nk, d.zrow, rbase[p] + cs[j], (uint32_t)(r / 32))] |= 1u << (r % 32);
        }
        // the key table's record of key kr (single class: records follow the header in key order)
        for (int w = 0; w < 4; w++) s.words[d.recw + 4 * r + w] = t.words[kClassHdr + 4 * kr + w];
    }
    s.lds_bytes = d.nwords * 4;
    return s.lds_bytes <= kMaskMaxIndexBytes;
}

namespace {

// Bulge families of every guide of q (see BulgeGuide).  Each compare key is
// emitted by exactly one variant, its first in expansion order, with the key's
// (first-wins) threshold; other variants of the same key carry kNeverHit.
// Leaves q.bulge empty when the query cannot use device-side expansion.
void encode_bulge(Query &q) {
    q.bulge.clear();
    const uint32_t D = q.D, R = q.R;
    if (D + R == 0 || D + R + 1 > (uint32_t)kMaxShifts) return;
    std::vector<char> emitted(q.keys.size(), 0);
    bool ok = true;
    for (const GuideLine &gl : q.guide_lines) {
        const std::string &g = gl.seq;
        if (g.size() > 64) { ok = false; break; }
        BulgeGuide bg;
        for (size_t k = 0; k < g.size(); k++) {
            const uint32_t mm = mismatch_classes(g[k]);
            if (!mm) continue;
            const uint64_t bit = 1ull << k;
            bg.care |= bit;
            for (int c = 0; c < 5; c++)
                if (mm & (1u << c)) bg.mask[c] |= bit;
            const char b = g[k];
            if (b == 'A' || b == 'C' || b == 'G' || b == 'T') {
                const uint32_t code = b == 'T' ? 0 : b == 'C' ? 1 : b == 'A' ? 2 : 3;
                if (code & 1) bg.lo |= bit;
                if (code & 2) bg.hi |= bit;
            } else {
                bg.concrete = false;
            }
        }
        std::vector<std::vector<uint32_t>> fam(D + R + 1);
        for (auto &t : bg.tmax) t = -1;
        bool any = false;
        expand_guide(g, D, R, q.rev, [&](const std::string &key, BulgeInfo &&, uint32_t u, uint32_t j) {
            if (fam[u].empty()) bg.j0[u] = (int32_t)j;
            else if ((int64_t)j != bg.j0[u] + (int64_t)fam[u].size() / 2) ok = false;  // splits must be consecutive
            auto it = std::lower_bound(q.keys.begin(), q.keys.end(), key,
                                       [](const CompareKey &a, const std::string &b) { return a.seq < b; });
            if (it == q.keys.end() || it->seq != key) { ok = false; return; }
            const size_t ki = (size_t)(it - q.keys.begin());
            uint32_t bias = kNeverHit;
            if (!emitted[ki]) {
                emitted[ki] = 1;
                bias = (uint32_t)(-(int64_t)it->threshold - 1);
                bg.tmax[u] = std::max<int32_t>(bg.tmax[u], it->threshold);
                any = true;
            }
            fam[u].push_back(bias);
            fam[u].push_back((uint32_t)ki);
        });
        if (!ok) break;
        if (!any) continue;  // every key of this guide is emitted by an earlier one
        for (uint32_t u = 0; u <= D + R; u++) {
            bg.n[u] = (uint32_t)(fam[u].size() / 2);
            bg.ent.insert(bg.ent.end(), fam[u].begin(), fam[u].end());
        }
        q.bulge.push_back(std::move(bg));
    }
    if (!ok) q.bulge.clear();
}

}  // namespace

// bits [0, x) of word w of a multi-word mask (x may be negative or past the end)
static uint32_t low_mask(int32_t x, int w) {
    const int32_t y = x - 32 * w;
    return y <= 0 ? 0u : y >= 32 ? ~0u : (1u << y) - 1u;
}

std::string unpadded_pattern(const Query &q) {
    return q.rev ? q.pattern.substr(0, q.P - q.D) : q.pattern.substr(q.D);
}

int build_bulge_table(const std::vector<BulgeRef> &qs, BulgeTable &t) {
    t = BulgeTable();
    if (qs.empty()) return 0;
    // The queries share the pattern without its padding and the PAM side;
    // their bulge sizes may differ.  The table's alignments are the union:
    // shift 0, DNA bulges 1..Dmax, RNA bulges 1..Rmax (shift Dmax + size);
    // the launch finds the windows of the query with the least padding
    // (Dmin) and each guide records its extra padding (bulge_sink).
    const Query &q0 = *qs[0].q;
    const std::string base = unpadded_pattern(q0);
    uint32_t Dmin = q0.D, Dmax = 0, Rmax = 0;
    const Query *qmin = &q0;
    uint64_t allcare = 0;
    bool gen = false;
    for (const BulgeRef &r : qs) {
        const Query &q = *r.q;
        if (q.D + q.R == 0 || q.rev != q0.rev || unpadded_pattern(q) != base) return 0;
        if (q.bulge.empty() && !q.keys.empty()) return 0;
        if (q.D < Dmin) Dmin = q.D, qmin = &q;
        Dmax = std::max(Dmax, q.D);
        Rmax = std::max(Rmax, q.R);
        for (const BulgeGuide &bg : q.bulge) {
            allcare |= bg.care;
            gen |= !bg.concrete;
        }
    }
    const int ns = (int)(Dmax + Rmax + 1);
    if (ns > kMaxShifts) return 0;
    const int gcb = allcare ? __builtin_ctzll(allcare) : 0;
    const int top = allcare ? 63 - __builtin_clzll(allcare) : 0;
    const int nw = top - gcb + 1 <= 32 ? 1 : 2;
    // window alignment of each shift: the left or right part of a variant
    // moves by the bulge size (cas-offinder.cpp:650-682 key layouts)
    const int sx = q0.rev ? 0 : (int)Dmin;
    for (int u = 0; u < ns; u++) {
        int delta = 0;
        if (u >= 1 && u <= (int)Dmax) delta = q0.rev ? u : -u;                    // DNA bulge of size u
        else if (u > (int)Dmax) delta = q0.rev ? -(u - (int)Dmax) : u - (int)Dmax;  // RNA bulge of size u - Dmax
        t.cb[u] = sx + delta + gcb;
        t.gap[u] = u > (int)Dmax ? (uint32_t)(u - (int)Dmax) : 0u;
        // every LDS word the kernel reads for a window start in the tile stays
        // inside the staged tile + halo, on both strands
        const int64_t P = qmin->P, halo = 32 * kHaloL;
        const int64_t plus_lo = t.cb[u] + halo, plus_hi = kTilePos - 1 + t.cb[u] + halo + 32 * (nw - 1);
        const int64_t minus_lo = P - 32 * nw - t.cb[u] + halo;
        const int64_t minus_hi = kTilePos - 1 + P - 32 * nw - t.cb[u] + halo + 32 * (nw - 1);
        if (plus_lo < 0 || minus_lo < 0 || (plus_hi >> 5) + 1 >= kLdsWords || (minus_hi >> 5) + 1 >= kLdsWords)
            return 0;
    }
    t.nw = nw;
    t.gen = gen;
    t.ns = ns;
    t.rev = q0.rev;
    t.pat = qmin->pat;
    t.dmin = Dmin;
    t.dpad = Dmax - Dmin;
    auto word = [&](uint64_t x, int w) { return (uint32_t)((x >> gcb) >> (32 * w)); };
    // care classes (one class of all-ones care for general tables)
    struct Item {
        const BulgeGuide *bg;
        uint32_t gbase;
        const Query *q;
    };
    std::map<std::vector<uint32_t>, std::vector<Item>> cls;
    std::vector<std::vector<uint32_t>> order;
    for (const BulgeRef &r : qs)
        for (const BulgeGuide &bg : r.q->bulge) {
            std::vector<uint32_t> c(nw, ~0u);
            if (!gen)
                for (int w = 0; w < nw; w++) c[w] = word(bg.care, w);
            auto it = cls.find(c);
            if (it == cls.end()) { order.push_back(c); cls[c].push_back({&bg, r.gbase, r.q}); }
            else it->second.push_back({&bg, r.gbase, r.q});
        }
    const int KP = bulge_planes(nw, gen), S = bulge_stride(nw, gen, ns);
    size_t rec_end = 0;
    for (auto &c : order) rec_end += kClassHdr + cls[c].size() * (size_t)S;
    std::vector<uint32_t> &out = t.words;
    out.reserve(rec_end * 2);
    out.resize(rec_end, 0);
    size_t p = 0;
    for (auto &c : order) {
        const auto &items = cls[c];
        for (int w = 0; w < nw; w++) out[p + w] = c[w];
        out[p + nw] = (uint32_t)items.size();
        p += kClassHdr;
        for (const Item &it : items) {
            const BulgeGuide &bg = *it.bg;
            // table shift u -> the guide's own family (its query's layout:
            // DNA sizes 1..D, then RNA sizes 1..R), -1 where it has none;
            // and where each own family's entries start in bg.ent
            const int Dq = (int)it.q->D, Rq = (int)it.q->R;
            int srcu[kMaxShifts];
            size_t eoff[kMaxShifts + 1] = {0};
            for (int v = 0; v <= Dq + Rq; v++) eoff[v + 1] = eoff[v] + 2 * (size_t)bg.n[v];
            for (int u = 0; u < ns; u++) {
                if (u == 0) srcu[u] = 0;
                else if (u <= (int)Dmax) srcu[u] = u <= Dq ? u : -1;
                else srcu[u] = u - (int)Dmax <= Rq ? Dq + (u - (int)Dmax) : -1;
            }
            uint32_t *rec = &out[p];
            if (gen) {
                for (int c5 = 0; c5 < 5; c5++)
                    for (int w = 0; w < nw; w++) rec[c5 * nw + w] = word(bg.mask[c5], w);
            } else {
                for (int w = 0; w < nw; w++) {
                    rec[w] = word(bg.lo, w);
                    rec[nw + w] = word(bg.hi, w);
                    rec[2 * nw + w] = word(bg.care, w);
                }
            }
            for (int u = 0; u < ns; u++) {
                const int v = srcu[u];
                rec[KP + 1 + u] = v < 0 || bg.tmax[v] < 0 ? kNeverHit
                                                          : (uint32_t)(-(int64_t)bg.tmax[v] - (int64_t)t.gap[u] - 1);
            }
            // variant block (rec is not used past this point: out grows):
            // [gbase, 0, 0, 0], then one family block per shift; the record
            // holds its offset and the guide's extra padding
            const size_t vb = out.size();
            if (vb > kVbOffsetMask) return 0;
            rec[KP] = (uint32_t)vb | (it.q->D - Dmin) << kVbPadShift;
            const int FB = bulge_family_words(nw), SW = bulge_seg_words(nw);
            out.resize(vb + 4 + (size_t)ns * FB, 0);
            out[vb] = it.gbase;
            const size_t e = out.size();
            for (int u = 0; u < ns; u++) {
                const int v = srcu[u];
                if (v < 0) continue;  // no family here: zero block, never passes its bias
                uint32_t *fb = &out[vb + 4 + (size_t)u * FB];
                const int32_t jb0 = bg.j0[v] - gcb;
                const uint32_t n = bg.n[v], nseg = std::min<uint32_t>(n, kBulgeSegs);
                fb[0] = (uint32_t)jb0;
                fb[1] = n;
                fb[2] = (uint32_t)(e + eoff[v]);
                fb[3] = nseg;
                // segment s holds splits [a, a + cnt); its lower bound counts
                // `left` below the first split, `left & right` inside and
                // `right` from the last split + gap on
                for (uint32_t sg = 0; sg < nseg; sg++) {
                    const uint32_t a = (uint32_t)((uint64_t)n * sg / nseg), b = (uint32_t)((uint64_t)n * (sg + 1) / nseg);
                    uint32_t *sr = fb + 4 + sg * SW;
                    sr[0] = a;
                    sr[1] = b - a;
                    const int32_t ja = jb0 + (int32_t)a, jz = jb0 + (int32_t)b - 1 + (int32_t)t.gap[u];
                    for (int w = 0; w < nw; w++) {
                        sr[2 + w] = low_mask(ja, w);
                        sr[2 + nw + w] = low_mask(jz, w) & ~low_mask(ja, w);
                        sr[2 + 2 * nw + w] = ~low_mask(jz, w);
                    }
                }
            }
            out.insert(out.end(), bg.ent.begin(), bg.ent.end());
            p += S;
        }
    }
    t.units = (uint32_t)order.size();
    for (const BulgeRef &r : qs) {
        t.nguides += r.q->bulge.size();
        t.nkeys += r.q->keys.size();
        t.care_sum += r.q->care_sum;
        for (const BulgeGuide &bg : r.q->bulge)
            for (int u = 0; u < ns; u++) t.tmax = std::max<int>(t.tmax, bg.tmax[u]);
    }
    return 1;
}

bool build_bulge_slices(const BulgeTable &t, const PatternDesc &pd, SliceIndex &s) {
    s = SliceIndex();
    if (t.gen || t.nw != 1 || t.units != 1 || t.words.size() < (size_t)kClassHdr) return false;
    const int ns = t.ns, KP = bulge_planes(1, false), S = bulge_stride(1, false, ns);
    const uint32_t care = t.words[0], n = t.words[1];
    if (n == 0 || n > kSliceMaxGuides || care == 0) return false;
    const int k0 = __builtin_ctz(care), kn = __builtin_popcount(care);
    if ((care >> k0) != (kn == 32 ? ~0u : (1u << kn) - 1)) return false;  // one run of care positions
    // filter families: window shift of each alignment relative to alignment 0
    SliceDesc &d = s.desc;
    int u_of[2 * 2] = {-1, -1, -1, -1};  // family f -> table shift u
    int dn = 0, dp = 0;
    for (int u = 1; u < ns; u++) {
        const int delta = t.cb[u] - t.cb[0];
        if (delta < -2 || delta > 2 || delta == 0) return false;
        if (delta < 0) dn = std::max(dn, -delta);
        else dp = std::max(dp, delta);
    }
    if (dn + dp != ns - 1 || ns > 3) return false;  // one alignment per shift, at most two families
    for (int u = 1; u < ns; u++) {
        const int delta = t.cb[u] - t.cb[0];
        const int f = delta < 0 ? -delta - 1 : dn + delta - 1;
        if (u_of[f] >= 0) return false;
        u_of[f] = u;
    }
    const int nf = dn + dp, off = std::max(dn, dp);
    // kernel width: the care run itself where a kernel of that width exists
    // (21, 23), else the next of 16 ... 32 (positions past the run masked)
    d.KN = kn == 21 || kn == 23 ? kn : kn <= 16 ? 16 : kn <= 20 ? 20 : kn <= 24 ? 24 : kn <= 28 ? 28 : 32;
    d.kn = (uint32_t)kn;
    d.dn = (uint32_t)dn;
    d.dp = (uint32_t)dp;
    d.fbase = t.cb[0] + k0 - off;
    d.care = care;
    d.nslots = n;
    d.nwords = (n + 31) / 32;
    // every LDS word the kernel reads for a window start in the tile stays
    // inside the staged tile + halo, on both strands (frame of fw words)
    const int fw = (int)d.KN + 2 * off <= 32 ? 1 : 2;
    {
        const int64_t P = pd.P, halo = 32 * kHaloL;
        const int64_t plus_lo = d.fbase + halo, plus_hi = kTilePos - 1 + d.fbase + halo + 32 * (fw - 1);
        const int64_t minus_lo = P - 32 * fw - d.fbase + halo;
        const int64_t minus_hi = kTilePos - 1 + P - 32 * fw - d.fbase + halo + 32 * (fw - 1);
        if (plus_lo < 0 || minus_lo < 0 || (plus_hi >> 5) + 1 >= kLdsWords || (minus_hi >> 5) + 1 >= kLdsWords)
            return false;
    }
    d.wstride = (uint32_t)((2 * d.KN + (kSliceTPlanes + 1) * nf + 1 + 3) & ~3u);
    d.slot_off = d.nwords * d.wstride;
    if (!slice_desc_valid(d, ns)) return false;  // e.g. care runs shorter than 13 positions
    // count bound of each filter family per guide record (-1: the family
    // never emits); the no-bulge key's threshold joins the guide's first
    // family (popc(M_0 & M_u) <= popc(M_0) for every u), family 0 when it
    // has none, so that a guide's families stay its own (a table may merge
    // queries of different bulge sizes)
    std::vector<std::array<int, 4>> bound(n);
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t *rec = &t.words[kClassHdr + (size_t)r * S];
        auto tof = [&](int u) {
            const uint32_t b = rec[KP + 1 + u];
            return b == kNeverHit ? -1 : (int)(-(int32_t)b - 1);
        };
        int first = 0;
        while (first < nf && tof(u_of[first]) < 0) first++;
        if (first == nf) first = 0;
        for (int f = 0; f < nf; f++) {
            int T = tof(u_of[f]);
            if (f == first) T = std::max(T, tof(0));
            bound[r][f] = std::min(T, (1 << kSliceTPlanes) - 1);
        }
    }
    auto top = [&](uint32_t r) { return *std::max_element(bound[r].begin(), bound[r].begin() + nf); };
    // slots grouped by the set of families their guides have (a table may
    // merge queries of different bulge sizes; the kernel skips a family no
    // guide of a word has), then by decreasing bound
    auto fams = [&](uint32_t r) {
        uint32_t m = 0;
        for (int f = 0; f < nf; f++) m |= (bound[r][f] >= 0 ? 1u : 0u) << f;
        return m;
    };
    // word layout: guides grouped by family set, each set ordered by bound
    // (either direction), the sets in some order, packed densely or each set
    // starting a new word (empty slots have no family and never pass); the
    // layout of the smallest estimated filter cost per candidate, in VALU
    // per word from the kernel's code: ~93 for the alignment-0 mismatches
    // and the queue, 13 more for the shared count of a word with a family
    // bounded on halves, and per family it evaluates 106 (whole split range)
    // or 140 (two halves).  Any layout gives the same records.
    std::vector<uint32_t> sets;  // family sets present, largest first
    for (uint32_t r = 0; r < n; r++)
        if (std::find(sets.begin(), sets.end(), fams(r)) == sets.end()) sets.push_back(fams(r));
    std::sort(sets.rbegin(), sets.rend());
    auto layout = [&](const std::vector<uint32_t> &ord, bool pure, std::vector<uint32_t> &slot) {
        uint32_t si = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (pure && i && fams(ord[i]) != fams(ord[i - 1])) si = (si + 31) & ~31u;
            slot[i] = si++;
        }
        return si;  // slots
    };
    auto cost = [&](const std::vector<uint32_t> &ord, const std::vector<uint32_t> &slot, uint32_t nsl) {
        const uint32_t nw = (nsl + 31) / 32;
        std::vector<uint32_t> wf(nw, 0), wfine(nw, 0);
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t w = slot[i] / 32, r = ord[i];
            wf[w] |= fams(r);
            for (int f = 0; f < nf; f++)
                if (bound[r][f] >= kSliceFineT) wfine[w] |= 1u << f;
        }
        uint64_t c = 0;
        for (uint32_t w = 0; w < nw; w++) {
            c += 93 + (wfine[w] ? 13 : 0);
            for (int f = 0; f < nf; f++)
                if ((wf[w] >> f) & 1) c += (wfine[w] >> f) & 1 ? 140 : 106;
        }
        return c;
    };
    std::vector<uint32_t> order, slot_of;
    uint32_t nsl = 0;
    uint64_t best = UINT64_MAX;
    std::vector<uint32_t> perm(sets.size());
    for (size_t i = 0; i < perm.size(); i++) perm[i] = (uint32_t)i;
    do {  // set orders (from largest-first) x bound directions x dense / padded
        for (uint32_t dirs = 0; dirs < (1u << sets.size()); dirs++) {
            auto rank = [&](uint32_t r) {
                for (size_t i = 0; i < perm.size(); i++)
                    if (sets[perm[i]] == fams(r)) return (uint32_t)i;
                return 0u;
            };
            std::vector<uint32_t> ord(n);
            for (uint32_t r = 0; r < n; r++) ord[r] = r;
            std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
                const uint32_t rx = rank(x), ry = rank(y);
                if (rx != ry) return rx < ry;
                return (dirs >> rx) & 1 ? top(x) < top(y) : top(x) > top(y);
            });
            for (const bool pure : {false, true}) {
                std::vector<uint32_t> slot(n);
                const uint32_t ns_ = layout(ord, pure, slot);
                if (ns_ > kSliceMaxGuides) continue;
                const uint64_t c = cost(ord, slot, ns_);
                if (c < best) {
                    best = c;
                    order = ord;
                    slot_of = slot;
                    nsl = ns_;
                }
            }
        }
    } while (std::next_permutation(perm.begin(), perm.end()));
    d.nslots = nsl;
    d.nwords = (nsl + 31) / 32;
    d.slot_off = d.nwords * d.wstride;
    if (!slice_desc_valid(d, ns)) return false;
    s.words.assign(d.slot_off + (size_t)kBulgeSlotWords * nsl, 0);
    for (uint32_t i = 0; i < n; i++) {  // slot slot_of[i] holds guide record order[i]
        const uint32_t si = slot_of[i], r = order[i];
        const uint32_t *rec = &t.words[kClassHdr + (size_t)r * S];
        uint32_t *blk = &s.words[(size_t)(si / 32) * d.wstride];
        const uint32_t bit = 1u << (si % 32);
        for (int k = 0; k < kn; k++) {
            if ((rec[0] >> (k0 + k)) & 1) blk[k] |= bit;
            if ((rec[1] >> (k0 + k)) & 1) blk[d.KN + k] |= bit;
        }
        for (int f = 0; f < nf; f++) {
            const int T = bound[r][f];
            if (T < 0) continue;
            uint32_t *fp = blk + 2 * d.KN + (kSliceTPlanes + 1) * f;
            for (int p = 0; p < kSliceTPlanes; p++)
                if ((T >> p) & 1) fp[p] |= bit;
            fp[kSliceTPlanes] |= bit;
        }
        for (int f = 0; f < nf; f++)
            if (bound[r][f] >= kSliceFineT) blk[2 * d.KN + (kSliceTPlanes + 1) * nf] |= 1u << f;
        uint32_t *slot = &s.words[d.slot_off + (size_t)kBulgeSlotWords * si];
        slot[0] = rec[0];
        slot[1] = rec[1];
        slot[2] = rec[KP];
        for (int u = 0; u < kMaxShifts; u++) slot[3 + u] = u < ns ? rec[KP + 1 + u] : kNeverHit;
    }
    s.words.resize((s.words.size() + 3) & ~(size_t)3, 0);
    return true;
}

bool build_bulge_index(const BulgeTable &t, MaskIndex &s) {
    s = MaskIndex();
    if (t.gen || t.nw != 1 || t.units != 1 || t.words.size() < (size_t)kClassHdr) return false;
    const int ns = t.ns, KP = bulge_planes(1, false), S = bulge_stride(1, false, ns);
    const uint32_t care = t.words[0], n = t.words[1];
    if (n == 0 || n > 128) return false;
    uint32_t brk = 1;  // ranges one variant can break: max(1, largest RNA gap)
    for (int u = 0; u < ns; u++) brk = std::max(brk, t.gap[u]);
    // per guide: the largest threshold among its emitting variants (-1: none)
    std::vector<int> tg(n, -1);
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t *rec = &t.words[kClassHdr + (size_t)r * S];
        for (int u = 0; u < ns; u++)
            if (rec[KP + 1 + u] != kNeverHit)
                tg[r] = std::max(tg[r], -(int32_t)rec[KP + 1 + u] - (int)t.gap[u] - 1);
    }
    const int tmax = *std::max_element(tg.begin(), tg.end());
    if (tmax < 0) return false;
    std::vector<int> pos;
    for (int b = 0; b < 32; b++)
        if ((care >> b) & 1) pos.push_back(b);
    const int C = (int)pos.size();
    MaskDesc &d = s.desc;
    int parts = tmax + (int)brk + 1;  // at least one exact range per surviving guide
    for (;; parts++) {
        if (parts > 8 || parts > C) return false;
        bool fits = true;
        for (int p = 0; p < parts && fits; p++) {
            const int g0 = (int)((int64_t)C * p / parts), g1 = (int)((int64_t)C * (p + 1) / parts);
            fits = g1 > g0 && pos[g1 - 1] + 1 - pos[g0] <= kMaskMaxSpan;
        }
        if (fits) break;
    }
    // as many ranges as fit, up to eight: each costs a row per alignment, but
    // every extra range raises the need by one and the survivors (whose
    // loop is the wave's longest) drop faster than the row reads grow
    for (int p2 = std::min(8, C); p2 > parts; p2--) {
        bool fits = true;
        for (int p = 0; p < p2 && fits; p++) {
            const int g0 = (int)((int64_t)C * p / p2), g1 = (int)((int64_t)C * (p + 1) / p2);
            fits = g1 > g0 && pos[g1 - 1] + 1 - pos[g0] <= kMaskMaxSpan;
        }
        if (fits) {
            parts = p2;
            break;
        }
    }
    d.parts = (uint32_t)parts;
    d.minex = 3;
    d.nrec = n;
    d.care = care;
    d.nk = n <= 32 ? 1 : n <= 64 ? 2 : 4;
    // slots in decreasing threshold order, so that a 32-guide word holds
    // guides of similar thresholds (its need follows its largest)
    std::vector<uint32_t> order(n);
    for (uint32_t r = 0; r < n; r++) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return tg[x] > tg[y]; });
    uint32_t rows = 0, rbase[kMaskMaxParts];
    for (int p = 0; p < parts; p++) {
        const int g0 = (int)((int64_t)C * p / parts), g1 = (int)((int64_t)C * (p + 1) / parts);
        const int a = pos[g0], span = pos[g1 - 1] + 1 - a;
        rbase[p] = rows;
        d.first[p] = (uint32_t)a | (uint32_t)span << 8 | rows << 16;
        rows += 1u << (2 * span);
    }
    d.zrow = rows++;
    for (int p = parts; p < kMaskMaxParts; p++) d.first[p] = d.zrow << 16;  // span 0: the zero row
    d.recw = (rows * d.nk + 3) & ~3u;
    // rows, then per slot [glo, ghi, variant block offset, filter bias per alignment]
    d.nwords = d.recw + kBulgeSlotWords * n;
    for (uint32_t w = 0; w < d.nk; w++) {
        int tw = -1;
        for (uint32_t r = 32 * w; r < std::min<uint32_t>(n, 32 * w + 32); r++) tw = std::max(tw, tg[order[r]]);
        // a word without an emitting guide keeps nothing
        const uint32_t need = tw < 0 ? 15u : (uint32_t)(parts - (int)brk - tw);
        d.need[w / 8] |= need << (4 * (w % 8));
    }
    s.words.assign(d.nwords, 0);
    for (uint32_t r = 0; r < n; r++) {  // slot r holds guide record order[r]
        const uint32_t *rec = &t.words[kClassHdr + (size_t)order[r] * S];
        uint32_t *slot = &s.words[d.recw + kBulgeSlotWords * r];
        slot[0] = rec[0];
        slot[1] = rec[1];
        slot[2] = rec[KP];
        for (int u = 0; u < kMaxShifts; u++) slot[3 + u] = u < ns ? rec[KP + 1 + u] : kNeverHit;
        const uint32_t glo = rec[0], ghi = rec[1];
        for (int p = 0; p < parts; p++) {
            const int a = (int)(d.first[p] & 0xFF), span = (int)((d.first[p] >> 8) & 0xFF);
            const uint32_t m = ((1u << span) - 1) & (care >> a);  // care bits of the range
            const uint32_t gl = (glo >> a) & m, gh = (ghi >> a) & m;
            for (uint32_t code = 0; code < (1u << (2 * span)); code++) {
                const uint32_t cl = code & ((1u << span) - 1), ch = code >> span;
                if ((cl & m) == gl && (ch & m) == gh) s.words[(size_t)(rbase[p] + code) * d.nk + r / 32] |= 1u << (r % 32);
            }
        }
    }
    s.lds_bytes = d.nwords * 4;
    return s.lds_bytes <= kBulgeIndexMaxBytes;
}

int encode_query(Query &q, std::string &err) {
    const uint32_t P = q.P;
    if (P == 0) { err = "Empty pattern."; return -1; }
    if (P > (uint32_t)kMaxPattern) {
        err = "Patterns longer than " + std::to_string(kMaxPattern) + " bases (with DNA-bulge slack) are not supported.";
        return -1;
    }
    // finder pattern: forward and reverse complement (cas-offinder.cpp:708-716)
    PatternDesc &pd = q.pat;
    memset(&pd, 0, sizeof(pd));
    pd.P = P;
    std::string rc = q.pattern;
    revcomp(rc);
    for (uint32_t k = 0; k < P; k++) {
        uint32_t mm = mismatch_classes(q.pattern[k]);
        if (mm) pd.f[pd.nf++] = k | (~mm & 31u) << 8;
        mm = mismatch_classes(rc[k]);
        if (mm) pd.r[pd.nr++] = k | (~mm & 31u) << 8;
    }
    std::vector<KeyRef> refs;
    for (size_t n = 0; n < q.keys.size(); n++) refs.push_back({&q.keys[n], (uint32_t)n});
    KeyTable t;
    if (build_key_table(refs, P, t, err)) return -1;
    q.key_words.swap(t.words);
    q.key_units = t.units;
    q.nw = t.nw;
    q.gen = t.gen;
    q.cbase = t.cbase;
    q.care_sum = t.care_sum;
    // one single-key table per key for the plain comparer (same offset and format)
    const int stride = key_stride(q.nw, q.gen);
    q.single_stride = q.gen ? stride : kClassHdr + stride;
    q.single_words.clear();
    for (size_t n = 0; n < q.keys.size(); n++) {  // at the query's offset and width
        uint32_t units;
        encode_records({refs[n]}, P, q.cbase, q.nw, q.gen, false, q.single_words, units);
    }
    encode_bulge(q);
    return 0;
}

}  // namespace cof
