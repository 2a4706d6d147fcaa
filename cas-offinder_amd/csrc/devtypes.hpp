// Types shared by the host engine and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace cof {

// Base classes as bit indices of a 5-bit class set.  Codes 0..3 are the UCSC
// 2bit codes (T=0, C=1, A=2, G=3); class 4 is any non-ACGT byte (N, ';', ...).
enum : uint32_t { kClsT = 0, kClsC = 1, kClsA = 2, kClsG = 3, kClsO = 4 };

constexpr int kMaxPattern = 64;   // longest padded pattern the kernels take
constexpr int kThreads = 256;     // workgroup size
constexpr int kWordsPerThread = 4;
constexpr int kTileWords = kThreads * kWordsPerThread;  // 32-position words per workgroup tile
constexpr int kTilePos = kTileWords * 32;
constexpr int kHaloL = 2;         // words staged before the tile
constexpr int kHaloR = 6;         // words staged after the tile
constexpr int kLdsWords = kTileWords + kHaloL + kHaloR;

// Finder pattern: the positions of the padded pattern (forward, and its
// reverse complement in forward text orientation) that can mismatch, each
// packed as position | (5-bit set of classes that MATCH there) << 8.  Entries
// are 32-bit so that kernels read them with scalar loads.
struct PatternDesc {
    uint32_t P;
    uint32_t nf, nr;
    uint32_t pad;
    uint32_t f[kMaxPattern];
    uint32_t r[kMaxPattern];
};

// Compare-key tables (bit b of a word <-> key position cbase + 32*word + b).
//  Concrete (every comparable key base is A/C/G/T): care classes, each a
//    kClassHdr-dword header [care[NW], nkeys, 0..] followed by nkeys records
//    [klo[NW], khi[NW], bias, key index] with klo/khi masked by care;
//    records are 4 dwords (NW = 1) or 8 (NW = 2).
//  General (IUPAC codes): records [mask[class * NW + word] for the classes
//    T, C, A, G, other (the positions where that class mismatches), bias,
//    key index]; 8 or 12 dwords.
//  bias = -(threshold + 1) as uint32: popcount(mismatches) + bias < 0 is a hit.
constexpr int kClassHdr = 4;
constexpr int key_stride(int nw, bool gen) { return gen ? (nw == 1 ? 8 : 12) : (nw == 1 ? 4 : 8); }
constexpr int key_bias_index(int nw, bool gen) { return (gen ? 5 : 2) * nw; }

// Bulge-family tables (device-side DNA/RNA-bulge expansion): a guide is scored
// at kMaxShifts window alignments at most (its own plus one per bulge family).
// Records, grouped by care class like concrete key tables ([care[NW], nguides,
// 0, 0] headers), are [guide planes][variant block word][bias per shift],
// padded to 4 dwords; planes are glo[NW], ghi[NW], care[NW] (concrete) or the
// five class masks [class * NW + word] (general).  The variant block word is
// the block's offset | the guide's extra padding << kVbPadShift (tables that
// merge bulge sizes).  A variant block is [gbase, 0, 0, 0]
// followed by one family block per shift: [first split (bit coordinate,
// int32), split count, entry offset, segment count] and kBulgeSegs segment
// records [first split index, split count, below[NW], inner[NW], above[NW]]
// (the masks of the segment's lower bound), then the entries [bias,
// query-local key index] of every split.
constexpr int kMaxShifts = 5;
constexpr int kBulgeSegs = 4;
constexpr int kVbPadShift = 29;
constexpr uint32_t kVbOffsetMask = (1u << kVbPadShift) - 1;
constexpr int bulge_planes(int nw, bool gen) { return gen ? 5 * nw : 3 * nw; }
constexpr int bulge_stride(int nw, bool gen, int ns) { return (bulge_planes(nw, gen) + 1 + ns + 3) & ~3; }
constexpr int bulge_seg_words(int nw) { return 2 + 3 * nw; }
constexpr int bulge_family_words(int nw) { return 4 + kBulgeSegs * bulge_seg_words(nw); }
constexpr uint32_t kNeverHit = 1u << 30;  // bias of a record or entry that never hits
// Bulge seed index (bulge_index_kernel): MaskDesc rows, then per slot
// [glo, ghi, variant block word, family filter bias of each alignment].
constexpr int kBulgeSlotWords = 3 + kMaxShifts;
constexpr uint32_t kBulgeIndexMaxBytes = 32 * 1024;  // static LDS + index stays below 64 KiB

// Bit-sliced bulge filter (bulge_slice_kernel): the guides of a single-class
// A/C/G/T bulge table with a contiguous care run of kn <= 32 positions, as
// 32-guide words.  Per word a block of wstride dwords: GL[KN], GH[KN] (bit g:
// the lo / hi code bit of slot 32 * word + g at the k-th care position), then
// per filter family f (dn window shifts -1..-dn, then dp shifts +1..+dp) six
// threshold planes T[0..5] and an active word: family f's count
// popc(M0 & Mf) over the care run must be at most T(g) for slot g to reach
// the exact counts; then a flags dword (bit f: family f's bounds in the word
// are large enough that it is bounded on two halves of its split range
// instead of the whole range; both are lower bounds of every split, so the
// choice only moves work between the filter and the exact counts).  Slots
// are ordered by decreasing bound, so
// that a word holds guides of similar thresholds.  After the blocks, per slot
// kBulgeSlotWords words [glo, ghi, variant block offset, family filter bias
// per alignment].
constexpr int kSliceTPlanes = 6;
constexpr int kSliceFineT = 6;  // smallest bound of a family bounded on halves
constexpr uint32_t kSliceMaxGuides = 256;
struct SliceDesc {
    uint32_t nslots, nwords;  // guides (slots) and 32-slot words
    uint32_t kn, KN;          // care positions, the kernel's position count (kn in (KN - 4, KN])
    uint32_t dn, dp;          // negative / positive window shifts of the filter families
    int32_t fbase;            // table-frame window position of frame bit 0
    uint32_t wstride;         // dwords per word block (a multiple of 4)
    uint32_t slot_off;        // first dword of the slot records
    uint32_t care;            // care mask of the class (table frame), for the exact counts
};
// What bulge_slice_kernel can run (the builder and the launcher both check
// it): at most two filter families, one per alignment of a table of ns, a
// care run of kn positions in (KN - 4, KN] for a kernel width KN of 16, 20,
// 24, 28 or 32, or exactly KN = kn for the widths 21 and 23 (the kernel
// masks the positions past a shorter run), word blocks holding the planes
// and bounds, slots after the blocks.
inline bool slice_desc_valid(const SliceDesc &sd, int ns) {
    const bool width = sd.KN == 16 || sd.KN == 20 || sd.KN == 24 || sd.KN == 28 || sd.KN == 32 ||
                       ((sd.KN == 21 || sd.KN == 23) && sd.kn == sd.KN);
    return ns >= 2 && ns <= 3 && sd.dn <= 2 && sd.dp <= 2 && (int)(sd.dn + sd.dp) == ns - 1 && width &&
           sd.kn <= sd.KN && sd.kn + 4 > sd.KN && sd.nslots >= 1 && sd.nslots <= kSliceMaxGuides &&
           sd.nwords == (sd.nslots + 31) / 32 && sd.wstride % 4 == 0 &&
           sd.wstride >= 2 * sd.KN + (kSliceTPlanes + 1) * (sd.dn + sd.dp) + 1 && sd.slot_off == sd.nwords * sd.wstride;
}

// Statistics counters updated once per workgroup are spread over kCountSlots
// slots kSlotStride words apart (one 128-byte line each): a single hot address
// serializes its atomics in one L2 channel.  The host sums the slots.
constexpr int kCountSlots = 64;
constexpr int kSlotStride = 16;

// Seed index of a single-class concrete key table (fused_seed_kernel): the
// comparable positions are split into `parts` = largest threshold + 1 bit
// ranges, so a window within threshold of a key matches at least one range of
// the key exactly (pigeonhole).  Per range, offs[obase + code .. + 1) delimit
// the ids of the records whose bases there (2-bit codes, lo bits then hi bits)
// equal `code`; a key with N inside a range is listed under every code it
// matches.
constexpr int kSeedMaxParts = 8;
struct SeedDesc {
    uint32_t parts;
    uint32_t nrec;                  // records copied to LDS (12 bytes each: klo, khi, bias)
    uint32_t noffs, nids;           // uint16 entries of the offset and id arrays
    uint32_t first[kSeedMaxParts];  // bit of the range's first position | length << 8
    uint32_t mask[kSeedMaxParts];   // the range's bits
    uint32_t obase[kSeedMaxParts];  // the range's first entry in offs
};

// Bit-set seed index of a single-class concrete key table of at most
// 32 * kMaskMaxWords keys (fused_mask_kernel).  The comparable positions are
// split into `parts` bit ranges of at most kMaskMaxSpan bits; per range and
// 2-bit code (lo bits then hi bits) a row of nk words (1, 2 or a multiple of
// 4; minex 1 only up to 8 words) has bit r set when
// record r's bases there equal the code (a key with N inside a range is set
// under every code).  A window within threshold t of a key matches at least
// parts - t of the key's ranges exactly; the kernel counts only the records
// that match at least `minex` ranges (parts >= largest threshold + minex;
// minex 3 below: a count per word).
// Layout (words, copied to LDS as is): rows, then at recw the records [klo,
// khi, bias, key index].  Row zrow is all zero: the unused parts up to the
// kernel's part count address it with span 0.  Rows of nk < 4 words lie one
// after another (row i at i * nk); wider rows are stored chunk-major, words
// [4c, 4c + 4) of row i at c * mask_plane(zrow) + 4 * i, so that the 16-byte
// chunk reads of lanes at random rows spread over all LDS banks (rows of 8,
// 16 or 24 consecutive words would start on a few banks only).
constexpr uint32_t mask_plane(uint32_t zrow) { return (zrow + 1) * 4; }
constexpr uint32_t mask_word(uint32_t nk, uint32_t zrow, uint32_t row, uint32_t w) {
    return nk < 4 ? row * nk + w : (w / 4) * mask_plane(zrow) + row * 4 + w % 4;
}
constexpr int kMaskMaxWords = 24;
constexpr int kMaskMaxSpan = 4;
// per-word counts (minex 3) of tables of kMaskWideWords or more words may use
// up to kMaskMaxParts ranges; the others at most kSeedMaxParts
constexpr int kMaskMaxParts = 12;
constexpr int kMaskWideWords = 12;
constexpr uint32_t kMaskMaxIndexBytes = 40 * 1024;
// minex 3: thresholds differ across the table; records are ordered by
// threshold and word w keeps the records that match at least need(w) =
// parts - (largest threshold in the word) ranges exactly (4-bit fields of
// `need`, word w at bits 4 * (w % 8) of need[w / 8]).
struct MaskDesc {
    uint32_t parts, nk, minex, nrec;
    uint32_t zrow, recw, nwords, care;  // nwords: rows and records, a multiple of 4
    uint32_t first[kMaskMaxParts];  // bit of the range's first position | span << 8 | first row << 16
    uint32_t need[(kMaskMaxWords + 7) / 8];
};

// Hit sort key: global compare-key index << kHitKeyShift | window start << 1 | strand.
constexpr int kHitKeyShift = 40;
constexpr uint64_t kMaxGenomePositions = 1ull << (kHitKeyShift - 1);
constexpr uint32_t kMaxSearchKeys = 1u << (64 - kHitKeyShift);

// Host-side hit record, identical to cof_hit of casoff.h.
struct DevHit {
    uint64_t pos;
    uint32_t key;
    uint16_t mm;
    uint8_t strand;
    uint8_t query;
};

struct DevGenome {
    const uint32_t *lo, *hi, *ot;  // planes, nwords + padding words
    const uint64_t *poison;        // sorted; poison[npoison] is a sentinel >= total + kMaxPattern
    const uint64_t *tile_ppos;     // per kTilePos tile: first poison position >= the tile start
    const uint32_t *tile_pidx;     // ... and its index in `poison`
    uint64_t npoison;
    uint64_t total;
    uint64_t nwords;
};

}  // namespace cof
