// Input parsing and compare-key expansion with the reference's semantics
// (cas-offinder.cpp:570-694), and the device encodings of pattern and keys.
#include "query.hpp"

#include <algorithm>
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
            // first / last non-N index as std::string::find_*_not_of returns them
            const uint64_t f = g.find_first_not_of('N'), l = g.find_last_not_of('N');
            if (D == 0) {
                add_key(m, g, id, thr, {"0", 0});
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
                        add_key(m, key, id, thr, {key == base ? "0" : std::to_string(i), idx});
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
                    add_key(m, key, id, thr, {bulge, rev ? -(int)j : (int)j});
                }
            }
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
    encode_records(keys, P, t.cbase, t.nw, t.gen, true, t.words, t.units);
    return 0;
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
        if (mm) { pd.fk[pd.nf] = (uint8_t)k; pd.fm[pd.nf] = (uint8_t)(~mm & 31u); pd.nf++; }
        mm = mismatch_classes(rc[k]);
        if (mm) { pd.rk[pd.nr] = (uint8_t)k; pd.rm[pd.nr] = (uint8_t)(~mm & 31u); pd.nr++; }
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
    return 0;
}

}  // namespace cof
