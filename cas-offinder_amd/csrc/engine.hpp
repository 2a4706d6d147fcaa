// Device context: resident genome, work lanes (stream + buffers), the
// autotuner and the two search paths (plain: reference launch model; batched:
// one fused launch per pattern group).
#pragma once

#include <hip/hip_runtime_api.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/casoff.h"
#include "devtypes.hpp"
#include "genome.hpp"
#include "query.hpp"

struct cof_query {
    std::shared_ptr<cof::Query> q;
};

namespace cof {
// Allocator whose value-less construct leaves trivial elements uninitialized, so
// resizing a hit vector does not zero-fill memory that is written right after.
// Blocks of 4 MiB or more are 2 MiB aligned and marked for transparent huge
// pages: decoding tens of millions of hits then faults in 2 MiB pages.
template <typename T>
struct UninitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <typename U>
    UninitAlloc(const UninitAlloc<U> &) {}
    static constexpr size_t kHuge = size_t(2) << 20;
    T *allocate(size_t n) {
        const size_t bytes = n * sizeof(T);
        if (bytes < 2 * kHuge) return std::allocator<T>::allocate(n);
        void *p = std::aligned_alloc(kHuge, (bytes + kHuge - 1) / kHuge * kHuge);
        if (!p) throw std::bad_alloc();
        (void)madvise(p, bytes, MADV_HUGEPAGE);
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n) {
        if (n * sizeof(T) < 2 * kHuge) std::allocator<T>::deallocate(p, n);
        else std::free(p);
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        if constexpr (sizeof...(A) == 0)
            ::new ((void *)p) U;
        else
            ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
}  // namespace cof

struct cof_result {
    // hits of query i: [qbegin[i], qend[i]), sorted by (key, pos, strand); the
    // queries' runs lie in launch order, not necessarily in query order
    std::vector<cof_hit, cof::UninitAlloc<cof_hit>> hits;
    std::vector<std::shared_ptr<cof::Query>> queries;
    std::vector<size_t> qbegin, qend;
    uint64_t genome_serial = 0;                     // the genome searched (cof_ctx::genome_serial)
};

namespace cof {

uint64_t finder_tiles(uint64_t slab_lo, uint64_t slab_hi);
hipError_t launch_finder_count(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                               uint32_t *tile_cnt, unsigned long long *cand_slots, hipStream_t s);
hipError_t scan_tile_counts(void *tmp, size_t *tmp_bytes, const uint32_t *cnt, uint32_t *incl, uint64_t n,
                            hipStream_t s);
hipError_t launch_finder_write(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                               const uint32_t *tile_incl, uint32_t *loci, uint8_t *flags, uint64_t cap,
                               hipStream_t s);
hipError_t launch_comparer(const DevGenome &g, uint32_t P, int nw, bool gen, uint32_t cbase, const uint32_t *table,
                           uint32_t gbase, uint64_t slab_lo, const uint32_t *loci, const uint8_t *flags, uint64_t n,
                           uint64_t *hkey, uint16_t *hmm, unsigned long long *count, uint64_t cap, hipStream_t s);
// variant: fused 0 = 32-key sign blocks, 1 = 8-key blocks, 2 = seed index
// (launch_seed; single-class concrete tables), 3 / 4 = bit-set seed index
// keeping records that match 1 / 2 ranges exactly (launch_mask; single-class
// concrete tables of at most 256 keys); bulge 0 = queued exact counts, 1 = in
// place, 2 = behind the bulge seed index (launch_bulge_index; single-class
// A/C/G/T tables of at most 128 guides).  tile_stride / max_tiles sample
// tiles (autotuner trials); 1 / 0 scan the whole genome.  Bulge variant 3
// runs behind the bit-sliced guide filter (launch_bulge_slice; single-class
// A/C/G/T tables of at most 256 guides, D + R <= 2).
constexpr int kFusedVariants = 5, kBulgeVariants = 4;
// One-time kernel attributes (dynamic LDS limits); called when a lane opens.
hipError_t prepare_kernels();
hipError_t launch_mask(const DevGenome &g, const PatternDesc &pd, const MaskDesc &md, uint32_t cbase,
                       const uint32_t *idx, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                       unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                       uint32_t max_tiles);
hipError_t launch_seed(const DevGenome &g, const PatternDesc &pd, const SeedDesc &sd, uint32_t lds_bytes,
                       uint32_t cbase, const uint32_t *table, const uint32_t *seed, uint64_t *hkey, uint16_t *hmm,
                       unsigned long long *count, unsigned long long *cand_count, uint64_t cap, hipStream_t s,
                       uint32_t tile_stride, uint32_t max_tiles);
hipError_t launch_fused(const DevGenome &g, const PatternDesc &pd, int nw, bool gen, uint32_t cbase,
                        const uint32_t *table, uint32_t nunits, uint32_t gbase, uint64_t *hkey, uint16_t *hmm,
                        unsigned long long *count, unsigned long long *cand_count, uint64_t cap, hipStream_t s,
                        int variant, uint32_t tile_stride, uint32_t max_tiles);
hipError_t launch_bulge(const DevGenome &g, const PatternDesc &pd, int nw, bool gen, int ns, const int32_t *cb,
                        const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, uint32_t nunits,
                        uint64_t *hkey,
                        uint16_t *hmm, unsigned long long *count, unsigned long long *cand_count, uint64_t cap,
                        hipStream_t s, int variant, uint32_t tile_stride, uint32_t max_tiles);
hipError_t launch_bulge_index(const DevGenome &g, const PatternDesc &pd, int ns, const int32_t *cb,
                              const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, const MaskDesc &md,
                              const uint32_t *idx, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                              unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                              uint32_t max_tiles);
hipError_t launch_bulge_slice(const DevGenome &g, const PatternDesc &pd, int ns, const int32_t *cb,
                              const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, const SliceDesc &sd,
                              const uint32_t *slices, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                              unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                              uint32_t max_tiles);
hipError_t sort_hits(void *tmp, size_t *tmp_bytes, const uint64_t *kin, uint64_t *kout, const uint16_t *vin,
                     uint16_t *vout, uint64_t n, unsigned end_bit, hipStream_t s);
// Packs n sorted hits (genome below 2^32 positions) into pos[i] (window
// start) and ms[i] (mismatches << 1 | strand), and koff[0..nkeys] into the
// first hit of each global key (nkeys: the search's key count).
hipError_t snap_count(const unsigned long long *count, unsigned long long *dst, hipStream_t s);
hipError_t pack_hits(const uint64_t *hkey, const uint16_t *hmm, uint64_t n, uint32_t nkeys, uint32_t *pos, uint8_t *ms,
                     uint32_t *koff, hipStream_t s);
// Kernel instance of the calling thread's last batched launch, as profilers
// name it ("cof::fused_mask_kernel<20, 3, 8, 2>"); empty after a reset when
// the launcher had nothing to launch.
const char *last_launched_kernel();
void reset_launched_kernel();

// One batched launch of a search: kernel instance, launch signature, variant,
// duration (HIP events around the launch on the lane's stream) and the
// (locus, strand) candidates it scanned.
struct LaunchRec {
    std::string kernel, sig;
    int variant = 0;
    double ms = 0;
    uint64_t candidates = 0;
};

// Pinned host buffer (hipHostMalloc), grow-only.
template <typename T>
struct HostBuf {
    T *p = nullptr;
    uint64_t n = 0;
    hipError_t reserve(uint64_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) { (void)hipHostFree(p); p = nullptr; n = 0; }
        hipError_t e = hipHostMalloc((void **)&p, std::max<uint64_t>(want, 1) * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = std::max<uint64_t>(want, 1);
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    uint64_t n = 0;  // capacity in elements
    hipError_t reserve(uint64_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc((void **)&p, std::max<uint64_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<uint64_t>(want, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Host side of a batched search: the key and bulge tables, seed indexes and
// launch signatures built from a query set, in one upload buffer.  Kept per
// lane for the query objects it was built from (it holds them) and reused
// when the next search names the same ones.
struct BatchPlan {
    uint64_t serial = 0;  // unique per plan built (the tables never change after)
    std::vector<std::shared_ptr<Query>> qs;
    std::vector<uint32_t> gbase;
    std::vector<KeyTable> ktab;
    std::vector<const Query *> krep;
    std::vector<SeedIndex> seeds;
    std::vector<char> has_seed;
    std::vector<MaskIndex> masks[2];
    std::vector<char> has_mask[2];
    std::vector<BulgeTable> btab;
    std::vector<MaskIndex> bidx;  // bulge seed indexes (has_bidx)
    std::vector<char> has_bidx;
    std::vector<SliceIndex> bslice;  // bit-sliced guide filters (has_bslice)
    std::vector<char> has_bslice;
    std::vector<size_t> bsoff;
    std::vector<uint32_t> kw;
    std::vector<size_t> koff, soff, boff, moff[2], bioff;
    std::vector<std::string> sig;
    std::vector<uint32_t> order;  // launch order: plan launch of each slot (most expected hits first)
};

// One in-order stream with its own work buffers; searches on different lanes
// may run concurrently over the same resident genome.
struct Lane {
    hipStream_t stream = nullptr;
    // collection stream: the sort and copy of each batched launch's hits run
    // there while the next launches run on `stream`
    hipStream_t stream2 = nullptr;
    HostBuf<unsigned long long> h_snap;  // hit counter after each batched launch
    std::vector<hipEvent_t> sev;         // after each launch's counter snapshot
    std::vector<hipEvent_t> cev2;        // per launch: before sort, after sort, after copy
    bool pipeline = true;                // per-launch collection (COF_PIPELINE=0: after the last launch)
    DevBuf<uint64_t> d_hkey, d_hkey2;  // hit sort keys (emitted, sorted)
    DevBuf<uint16_t> d_hmm, d_hmm2;    // hit mismatch counts (emitted, sorted)
    DevBuf<uint8_t> d_sort_tmp;
    HostBuf<uint64_t> h_hkey;
    HostBuf<uint16_t> h_hmm;
    // packed hits for the copy to the host (pack_hits)
    DevBuf<uint32_t> d_pos32, d_koff;
    DevBuf<uint8_t> d_ms;
    HostBuf<uint32_t> h_pos32, h_koff;
    HostBuf<uint8_t> h_ms;
    std::vector<hipEvent_t> cev;  // copy chunks
    // [0] hits; from kCandBase, per batched launch (or for the plain finder),
    // kCountSlots candidate-count slots kSlotStride words apart
    DevBuf<unsigned long long> d_count;
    DevBuf<uint32_t> d_loci;
    DevBuf<uint8_t> d_flags;
    DevBuf<uint32_t> d_tile_cnt, d_tile_incl;  // plain finder: entries per tile and their scan
    DevBuf<uint8_t> d_scan_tmp;
    DevBuf<uint32_t> d_keys;
    hipEvent_t ev[6] = {};
    std::vector<hipEvent_t> lev;       // around each batched launch
    cof_stats stats{};
    std::vector<LaunchRec> launches;   // batched launches of the last search
    // batched launch sequence of the last search, one HIP graph per launch
    // (events between them time each launch), replayed when a search has the
    // same tables, kernel arguments and buffers
    std::vector<hipGraphExec_t> graph;
    std::string graph_key;           // kernel arguments and buffers of the captured sequence
    uint64_t graph_plan = 0;         // serial of the plan the tables were last matched with
    std::vector<uint32_t> graph_kw;  // the tables the graphs read (uploaded before the capture)
    std::vector<int> graph_variants;
    std::vector<std::string> graph_kernels;
    void drop_graphs();
    bool graphs = true;
    bool trials = true;  // may run autotuner trials (off where latency matters: serving)
    bool pack_hits = true;  // packed hit copies where the genome allows (COF_HIT_PACK=0: 10-byte records)
    std::shared_ptr<BatchPlan> plan;  // of the last batched search
    hipError_t init();
    void release();
};

constexpr size_t kCandBase = kSlotStride;
constexpr size_t kCandWords = (size_t)kCountSlots * kSlotStride;  // per launch

// Autotuner of the batched kernels: the variant of each launch signature
// (kernel, word count, IUPAC, alignments, pattern, table size bucket, largest
// threshold) is chosen by timing every variant on a sample of tiles spread
// over the genome, the first time a context meets the signature, unless a
// decision for it was loaded from a tuner file (load / save).  Lanes that may
// not run trials (serving) take the decision of the nearest known signature:
// same fields but the last two, nearest size bucket, then nearest threshold.
// Shared by the lanes of a context.
struct Tuner {
    std::mutex mu;
    std::map<std::string, int> best;
    bool enabled = true;
    int force = -1;  // >= 0: every launch uses this variant, no trials
    bool get(const std::string &sig, int &v) {
        if (force >= 0) {
            v = force;
            return true;
        }
        std::lock_guard<std::mutex> lk(mu);
        auto it = best.find(sig);
        if (it == best.end()) return false;
        v = it->second;
        return true;
    }
    // exact decision, else the nearest known signature's; false when none shares the prefix
    bool get_near(const std::string &sig, int &v);
    void set(const std::string &sig, int v) {
        std::lock_guard<std::mutex> lk(mu);
        best[sig] = v;
    }
    // Tuner file: a JSON object {"<signature>": variant, ...}.  load merges
    // its decisions (COF_E_IO unreadable, COF_E_INVALID malformed, err set);
    // save writes every decision.
    int load(const std::string &path, std::string &err);
    int save(const std::string &path, std::string &err);
};

// Device buffers a caller brings to a batched search: key tables, launch
// counters (cleared by the caller), hit sort keys and mismatch counts.
struct ExtBufs {
    uint32_t *keys;
    size_t keys_words;
    unsigned long long *count;
    size_t count_words;
    uint64_t *hkey;
    uint16_t *hmm;
    uint64_t cap;
};
// Counter words a batched search of n queries needs (one launch per query at most).
inline size_t search_count_words(size_t n) { return kCandBase + n * kCandWords; }
// Batched search of `qs` into the caller's buffers: the tuner's decisions
// (nearest known signature, no trials), eager launches, no device or
// pinned allocation; *nhits = hits emitted (more than ext.cap: only the
// first cap were written, COF_E_NOMEM).  Hits are unsorted.
int run_search_noalloc(Lane &L, const DevGenome &dg, const std::vector<std::shared_ptr<Query>> &qs, Tuner *tuner,
                       const ExtBufs &ext, uint64_t &nhits, std::string &err);

// Sets up cof_search_noalloc's scratch array on `device` (the current
// device): called once per context open.
hipError_t noalloc_init(int device);

// Runs `qs` over the genome on lane L; *out gets sorted hits.  Returns a
// COF_* status with the message in err.  tuner may be null (default variants).
int run_search(Lane &L, const DevGenome &dg, uint64_t slab_positions, const std::vector<std::shared_ptr<Query>> &qs,
               int mode, Tuner *tuner, cof_result **out, std::string &err);

}  // namespace cof

struct cof_ctx {
    int device = 0;
    std::string err;
    cof::HostGenome genome;
    cof::DevBuf<uint32_t> d_lo, d_hi, d_ot;
    cof::DevBuf<uint64_t> d_poison, d_tile_ppos;
    cof::DevBuf<uint32_t> d_tile_pidx;
    cof::DevGenome dg{};
    bool genome_ready = false;
    // Identifies the loaded genome: unique over all contexts of the process,
    // renewed by every successful load or synth.  A result records the serial
    // it was searched on and formats only against that genome.
    uint64_t genome_serial = 0;
    std::atomic<int> servers{0};  // attached servers: genome reloads and close are refused meanwhile
    cof::Lane lane;                        // the context's own lane
    cof::Tuner tuner;
    uint64_t slab_positions = 1ull << 31;  // plain path window starts per finder launch
    double t_upload_ms = 0;                // duration of the last upload_genome

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    int upload_genome();
    // Makes g the context's genome (host copy, device planes, new serial).
    int install_genome(cof::HostGenome &&g);
    int search(const std::vector<std::shared_ptr<cof::Query>> &qs, int mode, cof_result **out);
};
