// Tuner decisions: nearest-signature lookup and the JSON tuner file.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "engine.hpp"

namespace cof {

namespace {

// signature = prefix | size bucket | largest threshold
bool split_sig(const std::string &sig, std::string &prefix, long &bucket, long &tmax) {
    const size_t b = sig.rfind('|');
    if (b == std::string::npos || b == 0) return false;
    const size_t a = sig.rfind('|', b - 1);
    if (a == std::string::npos) return false;
    prefix = sig.substr(0, a);
    char *end = nullptr;
    bucket = strtol(sig.c_str() + a + 1, &end, 10);
    if (end != sig.c_str() + b) return false;
    tmax = strtol(sig.c_str() + b + 1, &end, 10);
    return *end == 0;
}

}  // namespace

bool Tuner::get_near(const std::string &sig, int &v) {
    if (get(sig, v)) return true;
    std::string prefix, p2;
    long bk, tm, b2, t2;
    if (!split_sig(sig, prefix, bk, tm)) return false;
    std::lock_guard<std::mutex> lk(mu);
    long best_d = -1;
    for (auto it = best.lower_bound(prefix); it != best.end() && it->first.compare(0, prefix.size(), prefix) == 0;
         ++it) {
        if (!split_sig(it->first, p2, b2, t2) || p2 != prefix) continue;
        const long d = 1000 * std::labs(b2 - bk) + std::labs(t2 - tm);
        if (best_d < 0 || d < best_d) best_d = d, v = it->second;
    }
    return best_d >= 0;
}

int Tuner::load(const std::string &path, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot read tuner file " + path;
        return COF_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    std::map<std::string, int> got;
    size_t i = 0;
    auto ws = [&]() {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) i++;
    };
    auto bad = [&]() {
        err = "malformed tuner file " + path + " at byte " + std::to_string(i);
        return COF_E_INVALID;
    };
    ws();
    if (i >= t.size() || t[i++] != '{') return bad();
    ws();
    if (i < t.size() && t[i] == '}') i++;
    else
        for (;;) {
            ws();
            if (i >= t.size() || t[i++] != '"') return bad();
            const size_t e = t.find('"', i);
            if (e == std::string::npos) return bad();
            const std::string key = t.substr(i, e - i);
            if (key.find('\\') != std::string::npos) return bad();  // signatures need no escapes
            i = e + 1;
            ws();
            if (i >= t.size() || t[i++] != ':') return bad();
            ws();
            char *end = nullptr;
            const long v = strtol(t.c_str() + i, &end, 10);
            if (end == t.c_str() + i || v < 0 || v > 255) return bad();
            i = end - t.c_str();
            got[key] = (int)v;
            ws();
            if (i < t.size() && t[i] == ',') { i++; continue; }
            if (i < t.size() && t[i] == '}') { i++; break; }
            return bad();
        }
    ws();
    if (i != t.size()) return bad();
    std::lock_guard<std::mutex> lk(mu);
    for (auto &kv : got) best[kv.first] = kv.second;
    return COF_OK;
}

int Tuner::save(const std::string &path, std::string &err) {
    std::string text = "{\n";
    {
        std::lock_guard<std::mutex> lk(mu);
        size_t n = 0;
        for (auto &kv : best) {
            text += "  \"" + kv.first + "\": " + std::to_string(kv.second) + (++n < best.size() ? ",\n" : "\n");
        }
    }
    text += "}\n";
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write tuner file " + path;
        return COF_E_IO;
    }
    const bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
    if (fclose(f) != 0 || !ok) {
        err = "write failed: " + path;
        return COF_E_IO;
    }
    return COF_OK;
}

}  // namespace cof
