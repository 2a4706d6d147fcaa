// Record formatting (host): one TSV line per (hit, bulge info).
#include "format.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace cof {

std::string header_lines() {
    return "##Generated by Cas-OFFinder 3.0.0\n"
           "#Id\tBulge Type\tcrRNA\tDNA\tChromosome\tLocation\tDirection\tMismatches\tBulge Size\n";
}

namespace {

// std::string::substr with a clamped length; a start past the end yields "".
inline std::string sub(const std::string &s, size_t pos, size_t n = std::string::npos) {
    if (pos > s.size()) return std::string();
    return s.substr(pos, n);
}

inline std::string drop_tail(const std::string &s, size_t k) {
    // substr(0, size - k) where a wrapped length keeps the whole string
    return k <= s.size() ? s.substr(0, s.size() - k) : s;
}

inline bool all_digits(const std::string &s) {
    for (char c : s)
        if (c < '0' || c > '9') return false;
    return true;
}

}  // namespace

void format_hits(const HostGenome &g, const Query &q, const DevHit *hits, size_t n, std::string &out,
                 SummaryTable *summary) {
    const uint32_t P = q.P, D = q.D;
    std::string s(P, '\0'), rna, dna;
    char num[64];
    for (size_t h = 0; h < n; h++) {
        const DevHit &hit = hits[h];
        const CompareKey &key = q.keys[hit.key];
        g.text(hit.pos, P, &s[0]);
        const char dir = hit.strand ? '-' : '+';
        if (hit.strand) revcomp(s);
        for (uint32_t k = 0; k < P; k++)  // lowercase mismatched bases (cas-offinder.cpp:299-317)
            if (mismatch_classes(key.seq[k]) & class_of(s[k])) s[k] = (char)(s[k] + 32);
        const int f = g.file_of(hit.pos);
        const GenomeFile &gf = g.files[f < 0 ? 0 : f];
        const uint64_t L = hit.pos - gf.base;
        size_t idx = 0;  // last record starting at or before L (cas-offinder.cpp:400)
        {
            auto it = std::upper_bound(gf.chrpos.begin(), gf.chrpos.end(), L);
            idx = it == gf.chrpos.begin() ? 0 : (size_t)(it - gf.chrpos.begin()) - 1;
        }
        const std::string &name = gf.names.empty() ? std::string() : gf.names[idx];
        const uint64_t cpos = gf.chrpos.empty() ? 0 : gf.chrpos[idx];
        for (const BulgeInfo &bi : key.infos) {
            bool rv;
            uint32_t bidx;
            if (bi.second < 0) { rv = true; bidx = (uint32_t)(-bi.second); }
            else { rv = hit.strand != 0; bidx = (uint32_t)bi.second; }
            uint32_t bs;
            int off;
            const char *type;
            dna = s;
            if (all_digits(bi.first)) {  // DNA bulge or none
                bs = (uint32_t)atoi(bi.first.c_str());
                if (rv) {
                    off = 0;
                    rna = sub(key.seq, 0, bidx) + std::string(bs, '-') + sub(key.seq, bidx + bs);
                } else {
                    off = (int)D - (int)bs;
                    rna = sub(key.seq, off, bidx) + std::string(bs, '-') + sub(key.seq, off + bidx + bs);
                    dna = sub(dna, off);
                }
                if (bs == 0) {
                    type = "X";
                    if (rv) {
                        rna = drop_tail(rna, D);
                        dna = drop_tail(dna, D);
                    }
                } else {
                    type = "DNA";
                }
            } else {  // RNA bulge
                bs = (uint32_t)bi.first.size();
                if (rv) {
                    off = 0;
                    rna = sub(key.seq, 0, bidx) + bi.first + sub(key.seq, bidx);
                    dna = sub(dna, 0, bidx) + std::string(bs, '-') + sub(dna, bidx);
                    rna = drop_tail(rna, D + bs);
                    dna = drop_tail(dna, D + bs);
                } else {
                    off = (int)(D + bs);
                    rna = sub(key.seq, off, bidx) + bi.first + sub(key.seq, off + bidx);
                    dna = sub(dna, off, bidx) + std::string(bs, '-') + sub(dna, off + bidx);
                }
                type = "RNA";
            }
            out += key.id;
            out += '\t';
            out += type;
            out += '\t';
            out += rna;
            out += '\t';
            out += dna;
            out += '\t';
            out += name;
            out += '\t';
            snprintf(num, sizeof num, "%llu", (unsigned long long)(L - cpos + (uint64_t)(int64_t)off));
            out += num;
            out += '\t';
            out += dir;
            out += '\t';
            snprintf(num, sizeof num, "%u\t%u\n", (unsigned)hit.mm, bs);
            out += num;
            if (summary) {
                std::string k = key.id + "," + type + "," + std::to_string((int)bs) + "," + std::to_string((int)hit.mm);
                ++(*summary)[k];
            }
        }
    }
}

std::string summary_text(const SummaryTable &s) {
    static const char *hdr[] = {"Id", "Bulge Type", "Bulge Size", "Mismatches"};
    std::string out;
    for (auto &kv : s) {
        out += "##";
        const std::string &k = kv.first;
        size_t p = 0, cnt = 0;
        for (;;) {
            size_t c = k.find(',', p);
            const char *h = cnt < 4 ? hdr[cnt] : "?";
            out += h;
            out += '=';
            out += k.substr(p, c == std::string::npos ? std::string::npos : c - p);
            out += ';';
            cnt++;
            if (c == std::string::npos) break;
            p = c + 1;
        }
        out += "Number of Found Targets=" + std::to_string(kv.second) + "\n";
    }
    return out;
}

}  // namespace cof
