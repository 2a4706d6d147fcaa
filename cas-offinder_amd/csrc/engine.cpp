// Device context, work lanes and the two search paths.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <functional>
#include <thread>
#include <mutex>
#include <condition_variable>
#include <map>

using namespace cof;

namespace {

double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

#define HIP_TRY(expr, what)                                      \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) {                                  \
            err = std::string(what) + ": " + hipGetErrorString(e_); \
            return COF_E_DEVICE;                                 \
        }                                                        \
    } while (0)

// Key tables of all queries, concatenated; off[q] = first dword of query q.
// `single`: the per-key tables of the plain comparer instead of the fused table.
void key_tables(const std::vector<std::shared_ptr<Query>> &qs, std::vector<uint32_t> &kw, std::vector<size_t> &off,
                bool single) {
    off.resize(qs.size());
    kw.clear();
    for (size_t qi = 0; qi < qs.size(); qi++) {
        off[qi] = kw.size();
        const auto &t = single ? qs[qi]->single_words : qs[qi]->key_words;
        kw.insert(kw.end(), t.begin(), t.end());
    }
}

int search_plain(Lane &L, const DevGenome &dg, uint64_t slab_positions, const std::vector<std::shared_ptr<Query>> &qs,
                 const std::vector<uint32_t> &gbase, uint64_t &nhits, std::string &err) {
    double t_find = 0, t_cmp = 0;
    uint64_t ncand = 0, npairs = 0, nbase = 0, nmodel = 0, launches = 0, nloci = 0, npos = 0;
    std::vector<uint32_t> kw;
    std::vector<size_t> off;
    key_tables(qs, kw, off, true);
    // the per-key tables overwrite d_keys: the lane's captured batched graph
    // read the batched tables from it, so it must not be replayed without a
    // new upload
    L.graph_key.clear();
    L.graph_plan = 0;
    L.launches.clear();
    HIP_TRY(L.d_keys.reserve(kw.size() + 16), "hipMalloc(keys)");
    if (!kw.empty())
        HIP_TRY(hipMemcpyAsync(L.d_keys.p, kw.data(), kw.size() * 4, hipMemcpyHostToDevice, L.stream), "hipMemcpy(keys)");
    for (size_t qi = 0; qi < qs.size(); qi++) {
        const Query &q = *qs[qi];
        const uint64_t limit = dg.total >= q.P ? dg.total - q.P + 1 : 0;
        const int stride = q.single_stride;
        for (uint64_t lo = 0; lo < limit; lo += slab_positions) {
            const uint64_t hi = std::min(limit, lo + slab_positions);
            // finder pass 1: entries per tile, scanned into tile offsets
            const uint64_t tiles = finder_tiles(lo, hi);
            HIP_TRY(L.d_tile_cnt.reserve(tiles), "hipMalloc(finder tiles)");
            HIP_TRY(L.d_tile_incl.reserve(tiles), "hipMalloc(finder tiles)");
            size_t scan_bytes = 0;
            HIP_TRY(scan_tile_counts(nullptr, &scan_bytes, L.d_tile_cnt.p, L.d_tile_incl.p, tiles, L.stream), "scan size");
            HIP_TRY(L.d_scan_tmp.reserve(scan_bytes), "hipMalloc(scan)");
            unsigned long long *slots = L.d_count.p + kCandBase;
            HIP_TRY(hipMemsetAsync(slots, 0, kCandWords * sizeof(unsigned long long), L.stream), "hipMemsetAsync");
            HIP_TRY(hipEventRecord(L.ev[0], L.stream), "hipEventRecord");
            HIP_TRY(launch_finder_count(dg, q.pat, lo, hi, L.d_tile_cnt.p, slots, L.stream), "finder launch");
            HIP_TRY(scan_tile_counts(L.d_scan_tmp.p, &scan_bytes, L.d_tile_cnt.p, L.d_tile_incl.p, tiles, L.stream),
                    "finder scan");
            HIP_TRY(hipEventRecord(L.ev[1], L.stream), "hipEventRecord");
            uint32_t entries = 0;
            std::vector<unsigned long long> sl(kCandWords);
            HIP_TRY(hipMemcpyAsync(&entries, L.d_tile_incl.p + tiles - 1, sizeof entries, hipMemcpyDeviceToHost,
                                   L.stream),
                    "hipMemcpyAsync");
            HIP_TRY(hipMemcpyAsync(sl.data(), slots, kCandWords * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   L.stream),
                    "hipMemcpyAsync");
            HIP_TRY(hipStreamSynchronize(L.stream), "finder");
            unsigned long long found[2] = {entries, 0};
            for (int k = 0; k < kCountSlots; k++) found[1] += sl[(size_t)k * kSlotStride];
            // finder pass 2: entries at their offsets, in position order
            HIP_TRY(L.d_loci.reserve(found[0]), "hipMalloc(candidates)");
            HIP_TRY(L.d_flags.reserve(found[0]), "hipMalloc(candidates)");
            HIP_TRY(hipEventRecord(L.ev[4], L.stream), "hipEventRecord");
            HIP_TRY(launch_finder_write(dg, q.pat, lo, hi, L.d_tile_incl.p, L.d_loci.p, L.d_flags.p, L.d_loci.n,
                                        L.stream),
                    "finder launch");
            HIP_TRY(hipEventRecord(L.ev[5], L.stream), "hipEventRecord");
            launches += 3;
            ncand += found[1];
            nloci += found[0];
            npos += hi - lo;
            npairs += found[1] * q.keys.size();
            nbase += found[1] * q.care_sum;
            nmodel += found[1] * q.keys.size() * (uint64_t)q.nw * (q.gen ? 6 : 3);
            HIP_TRY(hipEventRecord(L.ev[2], L.stream), "hipEventRecord");
            for (size_t k = 0; k < q.keys.size(); k++) {
                HIP_TRY(launch_comparer(dg, q.P, q.nw, q.gen, (uint32_t)q.cbase, L.d_keys.p + off[qi] + k * (size_t)stride,
                                        gbase[qi], lo, L.d_loci.p, L.d_flags.p, found[0], L.d_hkey.p, L.d_hmm.p,
                                        L.d_count.p, L.d_hkey.n, L.stream),
                        "comparer launch");
                launches++;
            }
            HIP_TRY(hipEventRecord(L.ev[3], L.stream), "hipEventRecord");
            HIP_TRY(hipStreamSynchronize(L.stream), "comparer");
            t_find += ms_between(L.ev[0], L.ev[1]) + ms_between(L.ev[4], L.ev[5]);
            t_cmp += ms_between(L.ev[2], L.ev[3]);
        }
    }
    unsigned long long h = 0;
    HIP_TRY(hipMemcpyAsync(&h, L.d_count.p, sizeof h, hipMemcpyDeviceToHost, L.stream), "hipMemcpy(count)");
    HIP_TRY(hipStreamSynchronize(L.stream), "count");
    nhits = h;
    L.stats.t_finder_ms = t_find;
    L.stats.t_comparer_ms = t_cmp;
    L.stats.t_fused_ms = 0;
    L.stats.n_candidates = ncand;
    L.stats.n_pairs = npairs;
    L.stats.n_base_cmp = nbase;
    L.stats.n_model_ops = nmodel;
    L.stats.n_launches = launches;
    L.stats.t_tune_ms = 0;
    L.stats.graph = 0;
    L.stats.n_loci = nloci;
    L.stats.n_finder_positions = npos;
    return COF_OK;
}

// Probability that a uniform random base matches pattern code c (IUPAC:
// README.md:139-151).
static double base_match_prob(char c) {
    switch (c) {
        case 'A': case 'C': case 'G': case 'T': return 0.25;
        case 'R': case 'Y': case 'K': case 'M': case 'S': case 'W': return 0.5;
        case 'B': case 'D': case 'H': case 'V': return 0.75;
        default: return 1.0;
    }
}

// P(at most t of L bases mismatch), each independently with probability 3/4.
static double at_most_mismatches(int L, int t) {
    double term = std::pow(0.25, L), sum = 0;  // k = 0
    for (int k = 0; k <= std::min(t, L); k++) {
        sum += term;
        term *= (double)(L - k) / (k + 1) * 3.0;  // C(L, k+1) / C(L, k) x (3/4) / (1/4)
    }
    return std::min(sum, 1.0);
}

// Tables, indexes and launch signatures of a batched search over qs (see
// search_batched).
static int build_plan(const std::vector<std::shared_ptr<Query>> &qs, const std::vector<uint32_t> &gbase,
                      BatchPlan &P, std::string &err) {
    auto group_by = [&](const std::vector<size_t> &idx, auto keyf) {
        std::vector<std::vector<size_t>> groups;
        std::map<std::string, size_t> gi;
        for (size_t qi : idx) {
            const std::string k = keyf(*qs[qi]);
            auto it = gi.find(k);
            if (it == gi.end()) {
                gi[k] = groups.size();
                groups.push_back({qi});
            } else {
                groups[it->second].push_back(qi);
            }
        }
        return groups;
    };
    std::vector<size_t> bulged, plain;
    for (size_t qi = 0; qi < qs.size(); qi++) {
        if (qs[qi]->keys.empty()) continue;
        (qs[qi]->bulge.empty() ? plain : bulged).push_back(qi);
    }
    // bulge groups: the queries of one pattern (without its bulge padding)
    // and PAM side share a launch whatever their bulge sizes; where the
    // union of their alignments exceeds kMaxShifts or that table does not
    // qualify, one launch per (D, R)
    std::vector<BulgeTable> &btab = P.btab;
    std::vector<std::vector<size_t>> bq;  // queries of each bulge table
    auto dr_key = [](const Query &q) {
        return q.pattern + "|" + std::to_string(q.D) + "|" + std::to_string(q.R) + (q.rev ? "|r" : "|f");
    };
    for (auto &grp : group_by(bulged, [](const Query &q) { return unpadded_pattern(q) + (q.rev ? "|r" : "|f"); })) {
        uint32_t dmax = 0, rmax = 0;
        for (size_t qi : grp) dmax = std::max(dmax, qs[qi]->D), rmax = std::max(rmax, qs[qi]->R);
        const auto parts = group_by(grp, dr_key);
        if (parts.size() > 1 && dmax + rmax + 1 <= (uint32_t)kMaxShifts) {
            std::vector<BulgeRef> refs;
            for (size_t qi : grp) refs.push_back({qs[qi].get(), gbase[qi]});
            BulgeTable t;
            if (build_bulge_table(refs, t) == 1) {
                btab.push_back(std::move(t));
                bq.push_back(grp);
                continue;
            }
        }
        for (auto &part : parts) {
            std::vector<BulgeRef> refs;
            for (size_t qi : part) refs.push_back({qs[qi].get(), gbase[qi]});
            BulgeTable t;
            if (build_bulge_table(refs, t) == 1) {
                btab.push_back(std::move(t));
                bq.push_back(part);
            } else {
                plain.insert(plain.end(), part.begin(), part.end());
            }
        }
    }
    std::sort(plain.begin(), plain.end());
    const auto kgroups = group_by(plain, [](const Query &q) { return q.pattern; });
    std::vector<KeyTable> &ktab = P.ktab;
    std::vector<const Query *> &krep = P.krep;
    ktab.assign(kgroups.size(), KeyTable());
    krep.assign(kgroups.size(), nullptr);
    for (size_t g = 0; g < kgroups.size(); g++) {
        std::vector<KeyRef> refs;
        for (size_t qi : kgroups[g])
            for (size_t k = 0; k < qs[qi]->keys.size(); k++) refs.push_back({&qs[qi]->keys[k], gbase[qi] + (uint32_t)k});
        if (build_key_table(refs, qs[kgroups[g][0]]->P, ktab[g], err)) return COF_E_INVALID;
        krep[g] = qs[kgroups[g][0]].get();
    }
    // seed indexes of the key tables that qualify (an autotuner variant)
    auto &seeds = P.seeds;
    auto &has_seed = P.has_seed;
    auto &masks = P.masks;
    auto &has_mask = P.has_mask;
    seeds.assign(ktab.size(), SeedIndex());
    has_seed.assign(ktab.size(), 0);
    for (int x = 0; x < 2; x++) {
        masks[x].assign(ktab.size(), MaskIndex());
        has_mask[x].assign(ktab.size(), 0);
    }
    for (size_t g = 0; g < ktab.size(); g++) {
        std::vector<KeyRef> refs;
        for (size_t qi : kgroups[g])
            for (size_t k = 0; k < qs[qi]->keys.size(); k++) refs.push_back({&qs[qi]->keys[k], gbase[qi] + (uint32_t)k});
        has_seed[g] = build_seed_index(refs, krep[g]->P, ktab[g], seeds[g]);
        for (int x = 0; x < 2; x++) has_mask[x][g] = build_mask_index(refs, ktab[g], x + 1, masks[x][g]);
    }
    // one device buffer: key tables (each followed by its seed index), then bulge tables
    auto &kw = P.kw;
    auto &koff = P.koff, &soff = P.soff, &boff = P.boff;
    auto &moff = P.moff;
    koff.assign(ktab.size(), 0);
    soff.assign(ktab.size(), 0);
    boff.assign(btab.size(), 0);
    for (int x = 0; x < 2; x++) moff[x].assign(ktab.size(), 0);
    for (size_t g = 0; g < ktab.size(); g++) {
        koff[g] = kw.size();
        kw.insert(kw.end(), ktab[g].words.begin(), ktab[g].words.end());
        kw.resize((kw.size() + 3) & ~(size_t)3, 0);  // 16-byte aligned tables
        if (has_seed[g]) {
            soff[g] = kw.size();
            kw.insert(kw.end(), seeds[g].words.begin(), seeds[g].words.end());
            kw.resize((kw.size() + 3) & ~(size_t)3, 0);
        }
        for (int x = 0; x < 2; x++)
            if (has_mask[x][g]) {
                moff[x][g] = kw.size();
                kw.insert(kw.end(), masks[x][g].words.begin(), masks[x][g].words.end());
                kw.resize((kw.size() + 3) & ~(size_t)3, 0);
            }
    }
    auto &bidx = P.bidx;
    auto &has_bidx = P.has_bidx;
    auto &bioff = P.bioff;
    bidx.assign(btab.size(), MaskIndex());
    has_bidx.assign(btab.size(), 0);
    bioff.assign(btab.size(), 0);
    P.bslice.assign(btab.size(), SliceIndex());
    P.has_bslice.assign(btab.size(), 0);
    P.bsoff.assign(btab.size(), 0);
    for (size_t g = 0; g < btab.size(); g++) {
        boff[g] = kw.size();
        kw.insert(kw.end(), btab[g].words.begin(), btab[g].words.end());
        kw.resize((kw.size() + 3) & ~(size_t)3, 0);
        has_bidx[g] = build_bulge_index(btab[g], bidx[g]);
        if (has_bidx[g]) {
            bioff[g] = kw.size();
            kw.insert(kw.end(), bidx[g].words.begin(), bidx[g].words.end());
            kw.resize((kw.size() + 3) & ~(size_t)3, 0);
        }
        P.has_bslice[g] = build_bulge_slices(btab[g], btab[g].pat, P.bslice[g]);
        if (P.has_bslice[g]) {
            P.bsoff[g] = kw.size();
            kw.insert(kw.end(), P.bslice[g].words.begin(), P.bslice[g].words.end());
        }
    }
    const size_t nlaunch = ktab.size() + btab.size();
    auto bucket = [](uint64_t x) {
        int b = 0;
        while ((1ull << b) < x) b++;
        return std::to_string(b);
    };
    auto words = [](const void *p, size_t n) {
        std::string r;
        const uint32_t *w = static_cast<const uint32_t *>(p);
        for (size_t k = 0; k < n / 4; k++) r += std::to_string(w[k]) + ",";
        return r;
    };
    std::vector<std::string> &sig = P.sig;
    sig.assign(nlaunch, std::string());
    for (size_t i = 0; i < nlaunch; i++) {
        if (i < ktab.size()) {
            sig[i] = "fused|" + std::to_string(ktab[i].nw) + (ktab[i].gen ? "g|" : "c|") + krep[i]->pattern + "|" +
                     bucket(ktab[i].nkeys) + "|" + std::to_string(ktab[i].tmax);
        } else {
            const BulgeTable &t = btab[i - ktab.size()];
            sig[i] = "bulge|" + std::to_string(t.nw) + (t.gen ? "g|" : "c|") + std::to_string(t.ns) +
                     (t.rev ? "r|" : "f|") + words(t.cb, sizeof t.cb) + "|" + words(t.pat.f, 4 * t.pat.nf) + "|" +
                     bucket(t.nguides) + "|" + std::to_string(t.tmax);
        }
    }
    // launch order: most expected hits first.  Each launch's hits are
    // sorted, copied and decoded while the later launches run, so the last
    // launch's collection is what remains after the kernels: it should be
    // the one with the fewest hits.  Expected hits on a uniform random
    // genome: PAM match probability x the sum over compare keys of
    // P(at most t of the key's L bases mismatch), each with probability 3/4.
    std::vector<double> est(nlaunch, 0.0);
    auto add_query = [&](size_t i, const Query &q) {
        double pam = 1.0;
        for (char c : q.pattern) pam *= base_match_prob(c);
        double s = 0;
        for (const CompareKey &k : q.keys) {
            int L = 0;
            for (char c : k.seq) L += c == 'A' || c == 'C' || c == 'G' || c == 'T';
            s += at_most_mismatches(L, k.threshold);
        }
        est[i] += pam * s;
    };
    for (size_t g = 0; g < kgroups.size(); g++)
        for (size_t qi : kgroups[g]) add_query(g, *qs[qi]);
    for (size_t g = 0; g < bq.size(); g++)
        for (size_t qi : bq[g]) add_query(ktab.size() + g, *qs[qi]);
    P.order.resize(nlaunch);
    for (size_t i = 0; i < nlaunch; i++) P.order[i] = (uint32_t)i;
    std::stable_sort(P.order.begin(), P.order.end(), [&](uint32_t x, uint32_t y) { return est[x] > est[y]; });
    return COF_OK;
}

// Batched path, one fused launch (one genome pass) per group:
//  - queries with DNA/RNA bulges that share padded pattern, bulge sizes and
//    PAM side form a bulge group: bulge_kernel expands their variants on the
//    device from the guides;
//  - the other queries with the same padded pattern form a key group whose
//    compare keys go into one table (fused_kernel).
// Records carry global key indices.
// Per-launch collection (pipe): after each launch a one-thread kernel copies
// the hit counter into pinned memory, so the host learns where the launch's
// hits end.  The host keeps one launch ahead of the device: once launch i is
// done (launch i + 1 already queued), it queues the radix sort of launch i's
// range behind launch i + 1 on the lane's stream (the sort gets the whole
// GPU, between launches: on the second stream beside a search kernel it ran
// five times slower and only moved its time to the end of the search), then
// the range's copy on the second stream, and host threads decode copied
// ranges into the result at the same offsets while later launches run.  Every query's keys go to exactly one launch and
// every key of a launch belongs to its queries, so each query's hits are one
// sorted run inside its launch's range.  enqueue(i) queues launch i and its
// counter snapshot (event L.sev[i]).
static int collect_ranges(Lane &L, const std::vector<std::shared_ptr<Query>> &qs, const std::vector<uint32_t> &gbase,
                          size_t nlaunch, uint64_t cap, unsigned end_bit,
                          const std::function<hipError_t(size_t)> &enqueue, cof_result *&res, std::string &err) {
    res = nullptr;
    const size_t nq = qs.size();
    auto r = std::make_unique<cof_result>();
    r->hits.resize(cap);  // untouched beyond the hits found (no zero fill)
    r->qbegin.assign(nq, 0);
    r->qend.assign(nq, 0);
    std::vector<uint64_t> lo(nlaunch, 0), hi(nlaunch, 0);
    double t_decode = 0;
    auto decode = [&](size_t i) {
        const uint64_t b = lo[i], e = hi[i];
        if (e <= b) return;
        const auto td = std::chrono::steady_clock::now();
        const uint64_t *hk = L.h_hkey.p;
        const uint16_t *hm = L.h_hmm.p;
        // two host threads stay free: the one queueing launches and sorts,
        // and this decoder's own
        parallel_for(e - b, 1 << 16, [&](uint64_t x0, uint64_t x1) {
            x0 += b, x1 += b;
            size_t q = (size_t)(std::upper_bound(gbase.begin(), gbase.end() - 1, (uint32_t)(hk[x0] >> kHitKeyShift)) -
                                gbase.begin()) - 1;
            for (uint64_t x = x0; x < x1; x++) {
                const uint64_t k = hk[x];
                const uint32_t gk = (uint32_t)(k >> kHitKeyShift);
                while (q + 1 < nq && gk >= gbase[q + 1]) q++;  // keys are sorted
                cof_hit &h = r->hits[x];
                h.pos = (k >> 1) & (kMaxGenomePositions - 1);
                h.strand = (uint8_t)(k & 1);
                h.key = gk - gbase[q];
                h.mm = hm[x];
                h.query = (uint8_t)q;
            }
        }, std::max(1, n_host_threads() - 2));
        for (size_t q = 0; q < nq; q++) {
            const uint64_t *a = std::lower_bound(hk + b, hk + e, (uint64_t)gbase[q] << kHitKeyShift);
            const uint64_t *z = std::lower_bound(a, hk + e, (uint64_t)gbase[q + 1] << kHitKeyShift);
            if (z > a) r->qbegin[q] = (size_t)(a - hk), r->qend[q] = (size_t)(z - hk);
        }
        t_decode += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - td).count();
    };
    // decoder thread: ranges in launch order, each once its copy is done, so
    // that the loop below keeps queueing launches meanwhile
    std::mutex mu;
    std::condition_variable cv;
    size_t ready = 0;  // ranges whose copies are queued (lo / hi set)
    bool stop = false;
    hipError_t dec_err = hipSuccess;
    std::thread decoder([&] {
        for (size_t d = 0; d < nlaunch; d++) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return ready > d || stop; });
                if (ready <= d) return;
            }
            if (hi[d] > lo[d]) {
                const hipError_t x = hipEventSynchronize(L.cev2[3 * d + 2]);
                if (x != hipSuccess) {
                    dec_err = x;
                    return;
                }
            }
            decode(d);
        }
    });
    auto finish = [&](bool abort) {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = abort;
            if (!abort) ready = nlaunch;
        }
        cv.notify_all();
        decoder.join();
    };
    struct Joiner {  // error returns below: stop the decoder first
        std::function<void()> f;
        ~Joiner() { if (f) f(); }
    } joiner{[&] { finish(true); }};
    size_t tmp_bytes = L.d_sort_tmp.n;
    uint64_t prev = 0;
    bool over = false;
    auto t_last = std::chrono::steady_clock::now();
    if (nlaunch) HIP_TRY(enqueue(0), "batched launch");
    for (size_t i = 0; i < nlaunch; i++) {
        if (i + 1 < nlaunch) HIP_TRY(enqueue(i + 1), "batched launch");
        HIP_TRY(hipEventSynchronize(L.sev[i]), "batched launch");
        if (over) continue;  // the remaining launches still count their hits for the rerun
        const uint64_t e = reinterpret_cast<volatile unsigned long long *>(L.h_snap.p)[i];
        lo[i] = prev;
        hi[i] = std::max(e, prev);
        if (e > cap) {
            over = true;
            continue;
        }
        if (i + 1 == nlaunch) t_last = std::chrono::steady_clock::now();
        if (e > prev) {
            const uint64_t n = e - prev;
            HIP_TRY(hipEventRecord(L.cev2[3 * i], L.stream), "hipEventRecord");
            const uint64_t *sk = L.d_hkey.p;
            const uint16_t *sm = L.d_hmm.p;
            if (n > 1) {
                HIP_TRY(sort_hits(L.d_sort_tmp.p, &tmp_bytes, L.d_hkey.p + prev, L.d_hkey2.p + prev, L.d_hmm.p + prev,
                                  L.d_hmm2.p + prev, n, end_bit, L.stream),
                        "radix sort");
                sk = L.d_hkey2.p;
                sm = L.d_hmm2.p;
            }
            HIP_TRY(hipEventRecord(L.cev2[3 * i + 1], L.stream), "hipEventRecord");
            HIP_TRY(hipStreamWaitEvent(L.stream2, L.cev2[3 * i + 1], 0), "hipStreamWaitEvent");
            HIP_TRY(hipMemcpyAsync(L.h_hkey.p + prev, sk + prev, n * 8, hipMemcpyDeviceToHost, L.stream2),
                    "hipMemcpy(hits)");
            HIP_TRY(hipMemcpyAsync(L.h_hmm.p + prev, sm + prev, n * 2, hipMemcpyDeviceToHost, L.stream2),
                    "hipMemcpy(hits)");
            HIP_TRY(hipEventRecord(L.cev2[3 * i + 2], L.stream2), "hipEventRecord");
        }
        prev = hi[i];
        {
            std::lock_guard<std::mutex> lk(mu);
            ready = i + 1;
        }
        cv.notify_all();
    }
    joiner.f = nullptr;
    finish(over);
    if (over) {
        HIP_TRY(hipStreamSynchronize(L.stream2), "hit copy");
        return COF_OK;  // the caller grows the buffers and searches again
    }
    HIP_TRY(dec_err, "hit copy");
    double t_sort = 0, t_copy = 0;
    for (size_t i = 0; i < nlaunch; i++)
        if (hi[i] > lo[i]) {
            t_sort += ms_between(L.cev2[3 * i], L.cev2[3 * i + 1]);
            t_copy += ms_between(L.cev2[3 * i + 1], L.cev2[3 * i + 2]);
        }
    r->hits.resize(prev);
    r->queries = qs;
    L.stats.t_sort_ms = t_sort;
    L.stats.t_copy_ms = t_copy;
    L.stats.t_decode_ms = t_decode;
    L.stats.t_collect_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_last).count();
    res = r.release();
    return COF_OK;
}

int search_batched(Lane &L, const DevGenome &dg, const std::vector<std::shared_ptr<Query>> &qs,
                   const std::vector<uint32_t> &gbase, Tuner *tuner, uint64_t &nhits, std::string &err,
                   const ExtBufs *ext = nullptr, cof_result **pres = nullptr) {
    // the plan of the lane's previous search is reused for the same query
    // objects (it holds them, so their addresses identify them)
    L.stats.t_plan_ms = 0;
    if (!(L.plan && L.plan->qs == qs && L.plan->gbase == gbase)) {
        const auto tp0 = std::chrono::steady_clock::now();
        static std::atomic<uint64_t> next_serial{1};
        auto P = std::make_shared<BatchPlan>();
        if (const int rc = build_plan(qs, gbase, *P, err)) return rc;
        P->serial = next_serial.fetch_add(1);
        P->qs = qs;
        P->gbase = gbase;
        L.plan = std::move(P);
        L.stats.t_plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp0).count();
    }
    const BatchPlan &P = *L.plan;
    const auto &ktab = P.ktab;
    const auto &krep = P.krep;
    const auto &btab = P.btab;
    const auto &seeds = P.seeds;
    const auto &has_seed = P.has_seed;
    const auto &masks = P.masks;
    const auto &has_mask = P.has_mask;
    const auto &kw = P.kw;
    const auto &koff = P.koff, &soff = P.soff, &boff = P.boff;
    const auto &moff = P.moff;
    const auto &bidx = P.bidx;
    const auto &has_bidx = P.has_bidx;
    const auto &bioff = P.bioff;
    const auto &sig = P.sig;
    const auto &bslice = P.bslice;
    const auto &has_bslice = P.has_bslice;
    const auto &bsoff = P.bsoff;
    const auto &ord = P.order;
    const size_t nlaunch = ktab.size() + btab.size();
    // device buffers: the lane's, or the caller's (no allocation, no graphs,
    // no trials, no per-launch timing)
    uint32_t *d_keys;
    unsigned long long *d_count;
    size_t count_n;
    uint64_t *d_hkey;
    uint16_t *d_hmm;
    uint64_t cap;
    if (ext) {
        if (kw.size() > ext->keys_words) { err = "Key tables exceed the scratch buffer."; return COF_E_NOMEM; }
        d_keys = ext->keys;
        d_count = ext->count;
        count_n = ext->count_words;
        d_hkey = ext->hkey;
        d_hmm = ext->hmm;
        cap = ext->cap;
    } else {
        HIP_TRY(L.d_keys.reserve(kw.size() + 16), "hipMalloc(keys)");
        d_keys = L.d_keys.p;
        d_count = L.d_count.p;
        count_n = L.d_count.n;
        d_hkey = L.d_hkey.p;
        d_hmm = L.d_hmm.p;
        cap = L.d_hkey.n;
    }
    const bool graphs = L.graphs && !ext;
    // identity of the launch sequence: every kernel argument and the tables
    std::string gkey;
    if (graphs) {
        auto put = [&](const void *p, size_t n) { gkey.append(static_cast<const char *>(p), n); };
        const void *ptrs[] = {d_keys, d_hkey, d_hmm, d_count, dg.lo, dg.poison};
        put(ptrs, sizeof ptrs);
        put(&cap, sizeof cap);
        put(&dg.total, sizeof dg.total);
        for (size_t g = 0; g < ktab.size(); g++) {
            const int v[4] = {ktab[g].nw, ktab[g].gen, ktab[g].cbase, (int)ktab[g].units};
            put(v, sizeof v);
            put(&krep[g]->pat, sizeof(PatternDesc));
            put(&koff[g], sizeof koff[g]);
            put(&soff[g], sizeof soff[g]);
            put(&moff[0][g], sizeof moff[0][g]);
            put(&moff[1][g], sizeof moff[1][g]);
        }
        for (size_t g = 0; g < btab.size(); g++) {
            const BulgeTable &t = btab[g];
            const int v[5] = {t.nw, t.gen, t.ns, t.rev, (int)t.units};
            put(v, sizeof v);
            put(t.cb, sizeof t.cb);
            put(t.gap, sizeof t.gap);
            put(&t.pat, sizeof(PatternDesc));
            put(&t.dpad, sizeof t.dpad);
            put(&boff[g], sizeof boff[g]);
            put(&bioff[g], sizeof bioff[g]);
            put(&bsoff[g], sizeof bsoff[g]);
        }
        const size_t nkw = kw.size();
        put(&nkw, sizeof nkw);
        put(ord.data(), ord.size() * sizeof ord[0]);
    }
    // the tables the graphs read: those of the same plan object (its tables
    // never change), else compared word by word (new query objects with the
    // same tables still replay)
    const bool replay = graphs && !L.graph.empty() && gkey == L.graph_key &&
                        ((P.serial && L.graph_plan == P.serial) || L.graph_kw == kw);
    if (replay) L.graph_plan = P.serial;
    if (!replay && !kw.empty())
        HIP_TRY(hipMemcpyAsync(d_keys, kw.data(), kw.size() * 4, hipMemcpyHostToDevice, L.stream), "hipMemcpy(keys)");
    // counters: [0] hits, candidate slots per launch (run_search sized them for one launch per query)
    if (kCandBase + nlaunch * kCandWords > count_n) { err = "internal: counter buffer too small"; return COF_E_INVALID; }
    auto cand_slots = [&](size_t i) { return d_count + kCandBase + i * kCandWords; };
    // launch slot j runs plan launch i = ord[j]: i < ktab.size() is fused key
    // group i, else bulge group i - ktab.size(); counters, events, graphs and
    // variants are per slot
    auto launch = [&](size_t j, int v, uint32_t stride, uint32_t max_tiles) {
        const size_t i = ord[j];
        if (i < ktab.size() && (v == 3 || v == 4) && has_mask[v - 3][i])
            return launch_mask(dg, krep[i]->pat, masks[v - 3][i].desc, (uint32_t)ktab[i].cbase,
                               d_keys + moff[v - 3][i], d_hkey, d_hmm, d_count, cand_slots(j),
                               cap, L.stream, stride, max_tiles);
        if (i < ktab.size() && v == 2 && has_seed[i])
            return launch_seed(dg, krep[i]->pat, seeds[i].desc, seeds[i].lds_bytes, (uint32_t)ktab[i].cbase,
                               d_keys + koff[i], d_keys + soff[i], d_hkey, d_hmm, d_count,
                               cand_slots(j), cap, L.stream, stride, max_tiles);
        if (i < ktab.size())
            return launch_fused(dg, krep[i]->pat, ktab[i].nw, ktab[i].gen, (uint32_t)ktab[i].cbase,
                                d_keys + koff[i], ktab[i].units, 0, d_hkey, d_hmm, d_count,
                                cand_slots(j), cap, L.stream, v, stride, max_tiles);
        const size_t bg = i - ktab.size();
        const BulgeTable &t = btab[bg];
        if (v == 3 && has_bslice[bg])
            return launch_bulge_slice(dg, t.pat, t.ns, t.cb, t.gap, t.rev, t.dpad, d_keys + boff[bg], bslice[bg].desc,
                                      d_keys + bsoff[bg], d_hkey, d_hmm, d_count, cand_slots(j),
                                      cap, L.stream, stride, max_tiles);
        if (v == 2 && has_bidx[bg])
            return launch_bulge_index(dg, t.pat, t.ns, t.cb, t.gap, t.rev, t.dpad, d_keys + boff[bg], bidx[bg].desc,
                                      d_keys + bioff[bg], d_hkey, d_hmm, d_count, cand_slots(j),
                                      cap, L.stream, stride, max_tiles);
        return launch_bulge(dg, t.pat, t.nw, t.gen, t.ns, t.cb, t.gap, t.rev, t.dpad, d_keys + boff[bg],
                            t.units, d_hkey, d_hmm, d_count, cand_slots(j), cap, L.stream, v,
                            stride, max_tiles);
    };
    std::vector<int> variant(nlaunch, 0);
    // variants launch i can run (a cached choice for the same signature may
    // name an index the table does not qualify for: then the default runs)
    auto avail = [&](size_t j, int v) {
        const size_t i = ord[j];
        if (i >= ktab.size())
            return v == 0 || v == 1 || (v == 2 && has_bidx[i - ktab.size()]) || (v == 3 && has_bslice[i - ktab.size()]);
        switch (v) {
            case 0: case 1: return true;
            case 2: return (bool)has_seed[i];
            case 3: case 4: return (bool)has_mask[v - 3][i];
            default: return false;
        }
    };
    double t_tune = 0;
    if (replay) {
        variant = L.graph_variants;
    } else if (tuner && tuner->force >= 0) {
        for (size_t i = 0; i < nlaunch; i++)
            variant[i] = avail(i, tuner->force) ? tuner->force : ord[i] < ktab.size() ? 0 : 1;
    } else if (tuner && !(tuner->enabled && L.trials && !ext)) {
        for (size_t i = 0; i < nlaunch; i++)  // the nearest known decision, else the default
            if (!tuner->get_near(sig[ord[i]], variant[i]) || !avail(i, variant[i])) variant[i] = 0;
    } else if (tuner) {
        // time every variant of each new signature on tiles spread over the
        // genome; keep the default unless another is more than 3% faster
        constexpr uint32_t kTrialTiles = 8192;  // several waves of resident workgroups
        const uint64_t ntiles = (dg.total + kTilePos - 1) / kTilePos;
        const uint32_t stride = (uint32_t)std::max<uint64_t>(1, ntiles / kTrialTiles);
        bool tried = false;
        const auto tt0 = std::chrono::steady_clock::now();
        for (size_t i = 0; i < nlaunch; i++) {
            if (tuner->get(sig[ord[i]], variant[i]) && avail(i, variant[i])) continue;
            const int nv = ord[i] < ktab.size() ? kFusedVariants : kBulgeVariants;
            double best_ms = 0;
            int best = 0;
            for (int v = 0; v < nv; v++) {
                if (!avail(i, v)) continue;
                double ms = 1e30;
                for (int rep = 0; rep < 2; rep++) {
                    HIP_TRY(hipEventRecord(L.ev[2], L.stream), "hipEventRecord");
                    HIP_TRY(launch(i, v, stride, kTrialTiles), "trial launch");
                    HIP_TRY(hipEventRecord(L.ev[3], L.stream), "hipEventRecord");
                    HIP_TRY(hipEventSynchronize(L.ev[3]), "trial");
                    ms = std::min(ms, ms_between(L.ev[2], L.ev[3]));
                }
                if (v == 0) best_ms = ms;
                else if (ms < 0.97 * best_ms) best_ms = ms, best = v;
            }
            variant[i] = best;
            tuner->set(sig[ord[i]], best);
            tried = true;
        }
        if (tried)  // trial hits and candidate counts are discarded
            HIP_TRY(hipMemsetAsync(d_count, 0, count_n * sizeof(unsigned long long), L.stream), "hipMemsetAsync");
        t_tune = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tt0).count();
    }
    if (ext) {
        // eager launches and the hit counter back: no events, graphs or allocation
        for (size_t i = 0; i < nlaunch; i++)
            HIP_TRY(launch(i, variant[i], 1, 0), ord[i] < ktab.size() ? "fused launch" : "bulge launch");
        unsigned long long h = 0;
        HIP_TRY(hipMemcpyAsync(&h, d_count, sizeof h, hipMemcpyDeviceToHost, L.stream), "hipMemcpyAsync(count)");
        HIP_TRY(hipStreamSynchronize(L.stream), "fused");
        nhits = h;
        L.stats = cof_stats{};
        L.stats.n_launches = nlaunch;
        L.stats.n_hits = h;
        L.stats.genome_len = dg.total;
        L.launches.clear();
        return COF_OK;
    }
    const bool pipe = pres != nullptr;
    if (pres) *pres = nullptr;
    auto grow_events = [&](std::vector<hipEvent_t> &v, size_t n) -> hipError_t {
        while (v.size() < n) {
            hipEvent_t e = nullptr;
            const hipError_t x = hipEventCreate(&e);
            if (x != hipSuccess) return x;
            v.push_back(e);
        }
        return hipSuccess;
    };
    HIP_TRY(grow_events(L.lev, 2 * nlaunch), "hipEventCreate");
    unsigned end_bit = kHitKeyShift + 1;
    if (pipe) {
        // everything the collection needs is allocated before the launches
        HIP_TRY(grow_events(L.sev, nlaunch), "hipEventCreate");
        HIP_TRY(grow_events(L.cev2, 3 * nlaunch), "hipEventCreate");
        HIP_TRY(L.h_snap.reserve(nlaunch), "hipHostMalloc(counters)");
        HIP_TRY(L.d_hkey2.reserve(cap), "hipMalloc(sort)");
        HIP_TRY(L.d_hmm2.reserve(cap), "hipMalloc(sort)");
        HIP_TRY(L.h_hkey.reserve(cap), "hipHostMalloc(hits)");
        HIP_TRY(L.h_hmm.reserve(cap), "hipHostMalloc(hits)");
        unsigned keybits = 1;
        while ((1ull << keybits) < (uint64_t)gbase.back()) keybits++;
        end_bit = kHitKeyShift + keybits;
        size_t tmp_bytes = 0;
        HIP_TRY(sort_hits(nullptr, &tmp_bytes, L.d_hkey.p, L.d_hkey2.p, L.d_hmm.p, L.d_hmm2.p, cap, end_bit, L.stream2),
                "radix sort size");
        HIP_TRY(L.d_sort_tmp.reserve(tmp_bytes), "hipMalloc(sort)");
    }
    std::vector<std::string> kernels(nlaunch);
    if (!replay && graphs && nlaunch) {
        // capture each launch as a graph of its own (kernels only: counters
        // are cleared before, outside the graphs; the events between them
        // time each launch); a failed capture falls back to eager launches
        // for the rest of the lane's life
        L.drop_graphs();
        hipError_t e = hipSuccess;
        for (size_t i = 0; e == hipSuccess && i < nlaunch; i++) {
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            reset_launched_kernel();
            e = hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) break;
            e = launch(i, variant[i], 1, 0);
            kernels[i] = last_launched_kernel();
            const hipError_t e2 = hipStreamEndCapture(L.stream, &g);
            if (e == hipSuccess && e2 == hipSuccess && g) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            else if (e == hipSuccess) e = e2 != hipSuccess ? e2 : hipErrorUnknown;
            if (g) (void)hipGraphDestroy(g);
            if (e == hipSuccess) L.graph.push_back(x);
        }
        if (e == hipSuccess) {
            L.graph_key = gkey;
            L.graph_plan = P.serial;
            L.graph_kw = kw;
            L.graph_variants = variant;
            L.graph_kernels = kernels;
        } else {
            (void)hipGetLastError();
            L.drop_graphs();
            L.graphs = false;
        }
    }
    const bool use_graph = graphs && !L.graph.empty() && L.graph_key == gkey && P.serial && L.graph_plan == P.serial;
    L.stats.graph = use_graph ? (replay ? 2 : 1) : 0;
    HIP_TRY(hipEventRecord(L.ev[0], L.stream), "hipEventRecord");
    // launch i between its timing events (and, collecting per launch, the
    // hit counter snapshot after it)
    auto enqueue = [&](size_t i) -> hipError_t {
        hipError_t e = hipEventRecord(L.lev[2 * i], L.stream);
        if (e != hipSuccess) return e;
        if (use_graph) {
            e = hipGraphLaunch(L.graph[i], L.stream);
            kernels[i] = L.graph_kernels[i];
        } else {
            reset_launched_kernel();
            e = launch(i, variant[i], 1, 0);
            kernels[i] = last_launched_kernel();
        }
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(L.lev[2 * i + 1], L.stream)) != hipSuccess) return e;
        if (pipe) {
            if ((e = snap_count(d_count, L.h_snap.p + i, L.stream)) != hipSuccess) return e;
            e = hipEventRecord(L.sev[i], L.stream);
        }
        return e;
    };
    cof_result *res = nullptr;
    if (pipe) {
        const int rc = collect_ranges(L, qs, gbase, nlaunch, cap, end_bit, enqueue, res, err);
        if (rc) {
            (void)hipStreamSynchronize(L.stream);
            (void)hipStreamSynchronize(L.stream2);
            return rc;
        }
    } else {
        for (size_t i = 0; i < nlaunch; i++)
            HIP_TRY(enqueue(i), ord[i] < ktab.size() ? "fused launch" : "bulge launch");
    }
    HIP_TRY(hipEventRecord(L.ev[1], L.stream), "hipEventRecord");
    std::vector<unsigned long long> raw(kCandBase + nlaunch * kCandWords);
    HIP_TRY(hipMemcpyAsync(raw.data(), d_count, raw.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           L.stream),
            "hipMemcpyAsync(count)");
    HIP_TRY(hipStreamSynchronize(L.stream), "fused");
    std::vector<unsigned long long> cnt(2 + nlaunch, 0);  // [0] hits, [2 + launch] candidates
    cnt[0] = raw[0];
    for (size_t i = 0; i < nlaunch; i++)
        for (int sl = 0; sl < kCountSlots; sl++) cnt[2 + i] += raw[kCandBase + i * kCandWords + (size_t)sl * kSlotStride];
    L.stats.t_fused_ms = ms_between(L.ev[0], L.ev[1]);
    L.stats.t_comparer_ms = L.stats.t_fused_ms;
    L.stats.t_finder_ms = 0;
    L.launches.assign(nlaunch, LaunchRec());
    double launch_ms = 0;
    for (size_t i = 0; i < nlaunch; i++) {
        LaunchRec &lr = L.launches[i];
        lr.kernel = kernels[i];
        lr.sig = sig[ord[i]];
        lr.variant = variant[i];
        lr.ms = ms_between(L.lev[2 * i], L.lev[2 * i + 1]);
        lr.candidates = cnt[2 + i];
        launch_ms += lr.ms;
    }
    if (pipe) {  // the per-launch sorts between the launches count as collection (t_sort_ms)
        L.stats.t_fused_ms = launch_ms;
        L.stats.t_comparer_ms = launch_ms;
    }
    uint64_t ncand = 0, npairs = 0, nbase = 0, nmodel = 0;
    std::vector<size_t> slot(nlaunch);  // slot of each plan launch
    for (size_t j = 0; j < nlaunch; j++) slot[ord[j]] = j;
    for (size_t g = 0; g < ktab.size(); g++) {
        const uint64_t c = cnt[2 + slot[g]];
        ncand += c;
        npairs += c * ktab[g].nkeys;
        nbase += c * ktab[g].care_sum;
        nmodel += c * ktab[g].nkeys * (uint64_t)ktab[g].nw * (ktab[g].gen ? 6 : 3);
    }
    for (size_t g = 0; g < btab.size(); g++) {
        const BulgeTable &t = btab[g];
        const uint64_t c = cnt[2 + slot[ktab.size() + g]];
        ncand += c;
        npairs += c * t.nkeys;
        nbase += c * t.care_sum;
        // per (candidate, guide): masks at ns shifts, no-bulge popcount, one
        // AND + popcount per family, OR of the family scores
        const uint64_t ops = (uint64_t)t.nw * ((t.gen ? 5 : 2) * t.ns + 1 + 2 * (t.ns - 1)) + (t.ns - 1);
        nmodel += c * t.nguides * ops;
    }
    L.stats.n_candidates = ncand;
    L.stats.n_pairs = npairs;
    L.stats.n_base_cmp = nbase;
    L.stats.n_model_ops = nmodel;
    L.stats.n_launches = nlaunch;
    L.stats.t_tune_ms = t_tune;
    L.stats.n_loci = 0;
    L.stats.n_finder_positions = 0;
    nhits = cnt[0];
    if (res && nhits <= cap) *pres = res;
    else delete res;
    return COF_OK;
}

}  // namespace

hipError_t Lane::init() {
    hipError_t e = prepare_kernels();
    if (e != hipSuccess) return e;
    if ((e = hipStreamCreate(&stream)) != hipSuccess) return e;
    if ((e = hipStreamCreate(&stream2)) != hipSuccess) return e;
    if (const char *v = getenv("COF_PIPELINE")) pipeline = strcmp(v, "0") != 0;
    for (auto &x : ev)
        if ((e = hipEventCreate(&x)) != hipSuccess) return e;
    return hipSuccess;
}

void Lane::drop_graphs() {
    for (hipGraphExec_t x : graph) (void)hipGraphExecDestroy(x);
    graph.clear();
    graph_key.clear();
    graph_plan = 0;
    graph_kw.clear();
    graph_kernels.clear();
}

void Lane::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (hipEvent_t x : lev) (void)hipEventDestroy(x);
    lev.clear();
    plan.reset();
    d_hkey.release();
    d_hkey2.release();
    d_hmm.release();
    d_hmm2.release();
    d_sort_tmp.release();
    h_hkey.release();
    h_hmm.release();
    d_pos32.release();
    d_koff.release();
    d_ms.release();
    h_pos32.release();
    h_koff.release();
    h_ms.release();
    for (hipEvent_t x : cev) (void)hipEventDestroy(x);
    cev.clear();
    if (stream2) (void)hipStreamSynchronize(stream2);
    for (hipEvent_t x : sev) (void)hipEventDestroy(x);
    sev.clear();
    for (hipEvent_t x : cev2) (void)hipEventDestroy(x);
    cev2.clear();
    h_snap.release();
    if (stream2) (void)hipStreamDestroy(stream2);
    stream2 = nullptr;
    d_count.release();
    d_loci.release();
    d_flags.release();
    d_tile_cnt.release();
    d_tile_incl.release();
    d_scan_tmp.release();
    d_keys.release();
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x), x = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

int cof::run_search(Lane &L, const DevGenome &dg, uint64_t slab_positions, const std::vector<std::shared_ptr<Query>> &qs,
                    int mode, Tuner *tuner, cof_result **out, std::string &err) {
    *out = nullptr;
    if (qs.size() > 255) { err = "At most 255 queries per search."; return COF_E_INVALID; }
    if (mode != COF_MODE_PLAIN && mode != COF_MODE_BATCHED) { err = "Unknown search mode."; return COF_E_INVALID; }
    if (dg.total > kMaxGenomePositions) { err = "Genome too large for the hit key layout."; return COF_E_INVALID; }
    // global compare-key index base of every query
    std::vector<uint32_t> gbase(qs.size() + 1, 0);
    uint64_t nkeys = 0;
    for (size_t qi = 0; qi < qs.size(); qi++) {
        gbase[qi] = (uint32_t)nkeys;
        nkeys += qs[qi]->keys.size();
        if (nkeys >= kMaxSearchKeys) { err = "Too many compare keys in one search."; return COF_E_INVALID; }
    }
    gbase[qs.size()] = (uint32_t)nkeys;
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(L.d_count.reserve(kCandBase + qs.size() * kCandWords), "hipMalloc(counters)");
    if (!L.d_hkey.p) {
        HIP_TRY(L.d_hkey.reserve(1 << 20), "hipMalloc(hits)");
        HIP_TRY(L.d_hmm.reserve(1 << 20), "hipMalloc(hits)");
    }
    uint64_t nhits = 0;
    double tune_ms = 0;
    const bool pipe = mode == COF_MODE_BATCHED && L.pipeline;
    cof_result *pres = nullptr;
    for (int attempt = 0; attempt < 4; attempt++) {
        HIP_TRY(hipMemsetAsync(L.d_count.p, 0, L.d_count.n * sizeof(unsigned long long), L.stream), "hipMemsetAsync");
        int rc = mode == COF_MODE_PLAIN ? search_plain(L, dg, slab_positions, qs, gbase, nhits, err)
                                        : search_batched(L, dg, qs, gbase, tuner, nhits, err, nullptr,
                                                         pipe ? &pres : nullptr);
        if (rc) return rc;
        tune_ms += L.stats.t_tune_ms;
        if (nhits <= L.d_hkey.n) break;
        const uint64_t want = nhits + nhits / 4 + 1024;  // overflow: grow and rerun
        HIP_TRY(L.d_hkey.reserve(want), "hipMalloc(hits)");
        HIP_TRY(L.d_hmm.reserve(want), "hipMalloc(hits)");
    }
    if (nhits > L.d_hkey.n) { err = "Hit buffer overflow."; return COF_E_NOMEM; }
    if (pres) {  // collected launch by launch
        L.stats.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        L.stats.n_hits = nhits;
        L.stats.t_tune_ms = tune_ms;
        L.stats.genome_len = dg.total;
        *out = pres;
        return COF_OK;
    }
    const auto t_kern = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(L.ev[2], L.stream), "hipEventRecord");
    // sort on the device by (global key, start, strand), copy to pinned memory
    const uint64_t *hk = L.d_hkey.p;
    const uint16_t *hm = L.d_hmm.p;
    if (nhits > 1) {
        unsigned keybits = 1;
        while ((1ull << keybits) < (uint64_t)nkeys) keybits++;
        const unsigned end_bit = kHitKeyShift + keybits;
        size_t tmp_bytes = 0;
        HIP_TRY(sort_hits(nullptr, &tmp_bytes, L.d_hkey.p, L.d_hkey2.p, L.d_hmm.p, L.d_hmm2.p, nhits, end_bit, L.stream),
                "radix sort size");
        HIP_TRY(L.d_sort_tmp.reserve(tmp_bytes), "hipMalloc(sort)");
        HIP_TRY(L.d_hkey2.reserve(nhits), "hipMalloc(sort)");
        HIP_TRY(L.d_hmm2.reserve(nhits), "hipMalloc(sort)");
        HIP_TRY(sort_hits(L.d_sort_tmp.p, &tmp_bytes, L.d_hkey.p, L.d_hkey2.p, L.d_hmm.p, L.d_hmm2.p, nhits, end_bit,
                          L.stream),
                "radix sort");
        hk = L.d_hkey2.p;
        hm = L.d_hmm2.p;
    }
    auto *r = new cof_result();
    r->hits.resize(nhits);
    double t_decode = 0;
    if (nhits && L.pack_hits && dg.total < (1ull << 32) && nhits < (1ull << 32)) {
        // packed: 5 bytes per hit and the first hit of each key cross PCIe,
        // in chunks; each chunk is decoded by host threads while the next
        // one is copied
        const uint64_t nk = nkeys + 1;
        HIP_TRY(L.d_pos32.reserve(nhits), "hipMalloc(pack)");
        HIP_TRY(L.d_ms.reserve(nhits), "hipMalloc(pack)");
        HIP_TRY(L.d_koff.reserve(nk), "hipMalloc(pack)");
        HIP_TRY(L.h_pos32.reserve(nhits), "hipHostMalloc(pack)");
        HIP_TRY(L.h_ms.reserve(nhits), "hipHostMalloc(pack)");
        HIP_TRY(L.h_koff.reserve(nk), "hipHostMalloc(pack)");
        HIP_TRY(pack_hits(hk, hm, nhits, (uint32_t)nkeys, L.d_pos32.p, L.d_ms.p, L.d_koff.p, L.stream), "pack hits");
        HIP_TRY(hipEventRecord(L.ev[3], L.stream), "hipEventRecord");
        HIP_TRY(hipMemcpyAsync(L.h_koff.p, L.d_koff.p, nk * 4, hipMemcpyDeviceToHost, L.stream), "hipMemcpy(pack)");
        constexpr uint64_t kChunk = 1ull << 22;
        const uint64_t nch = (nhits + kChunk - 1) / kChunk;
        while (L.cev.size() < nch) {
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreate(&e), "hipEventCreate");
            L.cev.push_back(e);
        }
        for (uint64_t c = 0; c < nch; c++) {
            const uint64_t b = c * kChunk, n = std::min(nhits, b + kChunk) - b;
            HIP_TRY(hipMemcpyAsync(L.h_pos32.p + b, L.d_pos32.p + b, n * 4, hipMemcpyDeviceToHost, L.stream),
                    "hipMemcpy(pack)");
            HIP_TRY(hipMemcpyAsync(L.h_ms.p + b, L.d_ms.p + b, n, hipMemcpyDeviceToHost, L.stream), "hipMemcpy(pack)");
            HIP_TRY(hipEventRecord(L.cev[c], L.stream), "hipEventRecord");
        }
        HIP_TRY(hipEventRecord(L.ev[4], L.stream), "hipEventRecord");
        const uint32_t *koff = L.h_koff.p, *pos = L.h_pos32.p;
        const uint8_t *ms = L.h_ms.p;
        for (uint64_t c = 0; c < nch; c++) {
            HIP_TRY(hipEventSynchronize(L.cev[c]), "hit copy");
            const auto td = std::chrono::steady_clock::now();
            const uint64_t b0 = c * kChunk, e0 = std::min(nhits, b0 + kChunk);
            parallel_for(e0 - b0, 1 << 16, [&](uint64_t b, uint64_t e) {
                b += b0, e += b0;
                // the key of hit b: the last key whose first hit is at or before b
                uint64_t k = (uint64_t)(std::upper_bound(koff, koff + nk, (uint32_t)b) - koff) - 1;
                size_t q = (size_t)(std::upper_bound(gbase.begin(), gbase.end() - 1, (uint32_t)k) - gbase.begin()) - 1;
                for (uint64_t i = b; i < e; i++) {
                    while (koff[k + 1] <= i) k++;
                    while (q + 1 < qs.size() && k >= gbase[q + 1]) q++;
                    cof_hit &h = r->hits[i];
                    h.pos = pos[i];
                    h.key = (uint32_t)k - gbase[q];
                    h.mm = (uint16_t)(ms[i] >> 1);
                    h.strand = (uint8_t)(ms[i] & 1);
                    h.query = (uint8_t)q;
                }
            });
            t_decode += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - td).count();
        }
        L.stats.t_sort_ms = ms_between(L.ev[2], L.ev[3]);
        L.stats.t_copy_ms = ms_between(L.ev[3], L.ev[4]);
    } else {
        HIP_TRY(L.h_hkey.reserve(nhits), "hipHostMalloc(hits)");
        HIP_TRY(L.h_hmm.reserve(nhits), "hipHostMalloc(hits)");
        HIP_TRY(hipEventRecord(L.ev[3], L.stream), "hipEventRecord");
        if (nhits) {
            HIP_TRY(hipMemcpyAsync(L.h_hkey.p, hk, nhits * 8, hipMemcpyDeviceToHost, L.stream), "hipMemcpy(hits)");
            HIP_TRY(hipMemcpyAsync(L.h_hmm.p, hm, nhits * 2, hipMemcpyDeviceToHost, L.stream), "hipMemcpy(hits)");
        }
        HIP_TRY(hipEventRecord(L.ev[4], L.stream), "hipEventRecord");
        HIP_TRY(hipStreamSynchronize(L.stream), "hit copy");
        const auto td = std::chrono::steady_clock::now();
        const uint64_t *hkp = L.h_hkey.p;
        const uint16_t *hmp = L.h_hmm.p;
        parallel_for(nhits, 1 << 16, [&](uint64_t b, uint64_t e) {
            size_t q = 0;
            for (uint64_t i = b; i < e; i++) {
                const uint64_t k = hkp[i];
                const uint32_t gk = (uint32_t)(k >> kHitKeyShift);
                while (q + 1 < qs.size() && gk >= gbase[q + 1]) q++;  // keys are sorted
                cof_hit &h = r->hits[i];
                h.pos = (k >> 1) & (kMaxGenomePositions - 1);
                h.strand = (uint8_t)(k & 1);
                h.key = gk - gbase[q];
                h.mm = hmp[i];
                h.query = (uint8_t)q;
            }
        });
        t_decode = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - td).count();
        L.stats.t_sort_ms = ms_between(L.ev[2], L.ev[3]);
        L.stats.t_copy_ms = ms_between(L.ev[3], L.ev[4]);
    }
    L.stats.t_decode_ms = t_decode;
    r->queries = qs;
    r->qbegin.assign(qs.size(), 0);
    r->qend.assign(qs.size(), 0);
    for (size_t qi = 0; qi < qs.size(); qi++) {  // hits are sorted by query
        auto by_q = [](const cof_hit &x, size_t q) { return x.query < q; };
        r->qbegin[qi] = std::lower_bound(r->hits.begin(), r->hits.end(), qi, by_q) - r->hits.begin();
        r->qend[qi] = std::lower_bound(r->hits.begin(), r->hits.end(), qi + 1, by_q) - r->hits.begin();
    }
    const auto t1 = std::chrono::steady_clock::now();
    L.stats.t_total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    L.stats.t_collect_ms = std::chrono::duration<double, std::milli>(t1 - t_kern).count();
    L.stats.n_hits = nhits;
    L.stats.t_tune_ms = tune_ms;
    L.stats.genome_len = dg.total;
    *out = r;
    return COF_OK;
}

int cof::run_search_noalloc(Lane &L, const DevGenome &dg, const std::vector<std::shared_ptr<Query>> &qs,
                            Tuner *tuner, const ExtBufs &ext, uint64_t &nhits, std::string &err) {
    nhits = 0;
    if (qs.size() > 255) { err = "At most 255 queries per search."; return COF_E_INVALID; }
    if (dg.total > kMaxGenomePositions) { err = "Genome too large for the hit key layout."; return COF_E_INVALID; }
    if (ext.count_words < search_count_words(qs.size())) { err = "Counters exceed the scratch buffer."; return COF_E_NOMEM; }
    std::vector<uint32_t> gbase(qs.size() + 1, 0);
    uint64_t nkeys = 0;
    for (size_t qi = 0; qi < qs.size(); qi++) {
        gbase[qi] = (uint32_t)nkeys;
        nkeys += qs[qi]->keys.size();
        if (nkeys >= kMaxSearchKeys) { err = "Too many compare keys in one search."; return COF_E_INVALID; }
    }
    gbase[qs.size()] = (uint32_t)nkeys;
    const int rc = search_batched(L, dg, qs, gbase, tuner, nhits, err, &ext);
    if (rc) return rc;
    if (nhits > ext.cap) { err = "Hit buffer too small."; return COF_E_NOMEM; }
    return COF_OK;
}

// Host-to-device copy of a large pageable array through two pinned staging
// buffers: host threads fill one buffer while the DMA engine reads the
// other, so the transfer runs at the pinned-memory rate and overlaps the
// host-side copy.  done[b] marks the end of the last copy out of buffer b.
static hipError_t upload_staged(void *dst, const void *src, uint64_t bytes, HostBuf<uint8_t> (&stage)[2],
                                hipEvent_t (&done)[2], hipStream_t s) {
    const uint64_t chunk = stage[0].n;
    int b = 0;
    for (uint64_t off = 0; off < bytes; off += chunk, b ^= 1) {
        const uint64_t n = std::min(chunk, bytes - off);
        hipError_t e = hipEventSynchronize(done[b]);
        if (e != hipSuccess) return e;
        uint8_t *st = stage[b].p;
        const uint8_t *from = static_cast<const uint8_t *>(src) + off;
        parallel_for(n, 1 << 22, [&](uint64_t x0, uint64_t x1) { memcpy(st + x0, from + x0, x1 - x0); });
        if ((e = hipMemcpyAsync(static_cast<uint8_t *>(dst) + off, st, n, hipMemcpyHostToDevice, s)) != hipSuccess)
            return e;
        if ((e = hipEventRecord(done[b], s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

int cof_ctx::upload_genome() {
    genome_ready = false;
    const auto tu0 = std::chrono::steady_clock::now();
    const HostGenome &g = genome;
    const uint64_t nw = g.lo.size();  // includes the zero padding words
    HIP_TRY(d_lo.reserve(nw), "hipMalloc(genome lo)");
    HIP_TRY(d_hi.reserve(nw), "hipMalloc(genome hi)");
    HIP_TRY(d_ot.reserve(nw), "hipMalloc(genome ot)");
    {
        // the three bit planes through 2 x 64 MiB of pinned staging memory
        HostBuf<uint8_t> stage[2];
        hipEvent_t done[2] = {nullptr, nullptr};
        struct Free {
            HostBuf<uint8_t> (&st)[2];
            hipEvent_t (&ev)[2];
            ~Free() {
                for (auto &x : st) x.release();
                for (auto &e : ev)
                    if (e) (void)hipEventDestroy(e);
            }
        } free_on_exit{stage, done};
        const uint64_t chunk = std::min<uint64_t>(64ull << 20, std::max<uint64_t>(nw * 4, 1));
        for (int b = 0; b < 2; b++) {
            HIP_TRY(stage[b].reserve(chunk), "hipHostMalloc(genome staging)");
            HIP_TRY(hipEventCreate(&done[b]), "hipEventCreate");
            HIP_TRY(hipEventRecord(done[b], lane.stream), "hipEventRecord");
        }
        HIP_TRY(upload_staged(d_lo.p, g.lo.data(), nw * 4, stage, done, lane.stream), "hipMemcpy(genome)");
        HIP_TRY(upload_staged(d_hi.p, g.hi.data(), nw * 4, stage, done, lane.stream), "hipMemcpy(genome)");
        HIP_TRY(upload_staged(d_ot.p, g.ot.data(), nw * 4, stage, done, lane.stream), "hipMemcpy(genome)");
        HIP_TRY(hipStreamSynchronize(lane.stream), "hipMemcpy(genome)");
    }
    std::vector<uint64_t> pz(g.poison);
    pz.push_back(~0ull);  // sentinel
    HIP_TRY(d_poison.reserve(pz.size()), "hipMalloc(poison)");
    HIP_TRY(hipMemcpy(d_poison.p, pz.data(), pz.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(poison)");
    // per tile: first poison at or after the tile start, and its index
    const uint64_t ntiles = g.total / kTilePos + 2;
    std::vector<uint64_t> tpos(ntiles);
    std::vector<uint32_t> tidx(ntiles);
    for (uint64_t t = 0; t < ntiles; t++) {
        const size_t i = std::lower_bound(g.poison.begin(), g.poison.end(), t * (uint64_t)kTilePos) - g.poison.begin();
        tidx[t] = (uint32_t)i;
        tpos[t] = pz[i];
    }
    HIP_TRY(d_tile_ppos.reserve(ntiles), "hipMalloc(tiles)");
    HIP_TRY(d_tile_pidx.reserve(ntiles), "hipMalloc(tiles)");
    HIP_TRY(hipMemcpy(d_tile_ppos.p, tpos.data(), ntiles * 8, hipMemcpyHostToDevice), "hipMemcpy(tiles)");
    HIP_TRY(hipMemcpy(d_tile_pidx.p, tidx.data(), ntiles * 4, hipMemcpyHostToDevice), "hipMemcpy(tiles)");
    dg.tile_ppos = d_tile_ppos.p;
    dg.tile_pidx = d_tile_pidx.p;
    dg.lo = d_lo.p;
    dg.hi = d_hi.p;
    dg.ot = d_ot.p;
    dg.poison = d_poison.p;
    dg.npoison = g.poison.size();
    dg.total = g.total;
    dg.nwords = g.nwords();
    genome_ready = true;
    t_upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tu0).count();
    return COF_OK;
}

int cof_ctx::install_genome(HostGenome &&g) {
    static std::atomic<uint64_t> serials{0};
    genome = std::move(g);
    genome_serial = ++serials;
    return device < 0 ? COF_OK : upload_genome();
}

int cof_ctx::search(const std::vector<std::shared_ptr<Query>> &qs, int mode, cof_result **out) {
    *out = nullptr;
    if (!genome_ready) return fail(COF_E_INVALID, "No genome loaded.");
    const int rc = run_search(lane, dg, slab_positions, qs, mode, &tuner, out, err);
    if (rc == COF_OK) (*out)->genome_serial = genome_serial;
    return rc;
}
