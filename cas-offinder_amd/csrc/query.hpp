// Query = one reference input file: genome path, padded pattern with DNA/RNA
// bulge sizes and the compare keys expanded from the guides
// (cas-offinder.cpp:590-694), plus the device encodings of pattern and keys.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "devtypes.hpp"

namespace cof {

struct BulgeInfo {
    std::string first;  // "0", the DNA bulge size, or the RNA-bulge bases
    int second;         // bulge index, negated for a reversed PAM
};

struct CompareKey {
    std::string seq;
    std::string id;
    uint16_t threshold = 0;
    std::vector<BulgeInfo> infos;
};

struct GuideLine {
    std::string seq;  // uppercase
    std::string id;
    uint16_t threshold = 0;
};

// A guide with every DNA/RNA-bulge variant it expands to, for device-side
// expansion.  Shift 0 is the guide's own alignment (the no-bulge key when D = 0);
// shift i (1..D) is the family of DNA bulges of size i and shift D + i
// (1..R) the family of RNA bulges of size i.  A variant with split j counts
//   mm(j) = popcount(left & bits[0, j)) + popcount(right & ~bits[0, j + gap))
// over the mismatch masks of the guide at two alignments (bit k <-> guide
// position k): the left part at the family's shift and the right part at
// shift 0 for a forward PAM, the other way round for a reversed PAM; gap is the
// RNA-bulge length (0 for DNA bulges).
struct BulgeGuide {
    uint64_t care = 0, lo = 0, hi = 0;  // positions that can mismatch; 2-bit code planes there
    uint64_t mask[5] = {};              // per class: positions where that class mismatches
    bool concrete = true;               // every care position is A/C/G/T
    int32_t j0[kMaxShifts] = {};        // first split of each family (guide position)
    uint32_t n[kMaxShifts] = {};        // splits of each family
    int32_t tmax[kMaxShifts] = {};      // largest threshold among emitting variants, -1 if none
    std::vector<uint32_t> ent;          // per family, per split: bias, query-local key index
};

struct Query {
    std::string chrdir;
    std::string pattern;  // padded, uppercase
    uint32_t P = 0, D = 0, R = 0;
    bool rev = false;                   // reversed (5') PAM: pattern line does not start with N
    std::vector<CompareKey> keys;  // bytewise key order (std::map order)
    std::vector<GuideLine> guide_lines;
    std::vector<BulgeGuide> bulge;      // bulge families of the guides (D + R > 0, D + R < kMaxShifts)

    // device encodings (encode_query)
    PatternDesc pat{};
    int nw = 1;
    bool gen = false;  // some key has an IUPAC code at a comparable position
    int cbase = 0;
    std::vector<uint32_t> key_words;     // fused-kernel table: key_units care classes / records
    uint32_t key_units = 0;
    std::vector<uint32_t> single_words;  // one single-key table per key (plain comparer)
    int single_stride = 0;               // dwords per single-key table
    uint64_t guides = 0;              // guide lines read
    uint64_t care_sum = 0;            // sum over keys of positions that can mismatch
};

// 5-bit set of base classes that count as a mismatch against pattern/key char
// `p` (bit kClsT..kClsO).  IUPAC (README.md:139-151): a concrete base
// mismatches every other class including non-ACGT bytes; a degenerate code
// mismatches exactly the concrete bases outside its set; N and any other byte
// never mismatch.
uint32_t mismatch_classes(char p);
uint32_t class_of(char genome_char);
void revcomp(std::string &s);  // cas-offinder.cpp:35-58 complement table, then reverse

int parse_query(const char *text, size_t len, Query &q, std::string &err);
int encode_query(Query &q, std::string &err);

// Fused-kernel key table of any set of compare keys of one padded pattern
// length P: common key offset cbase, care classes merged across the keys,
// every record carrying the caller's key index.
struct KeyTable {
    std::vector<uint32_t> words;
    uint32_t units = 0;  // care classes (concrete) or padded record count (general)
    int nw = 1;
    bool gen = false;
    int cbase = 0;
    uint64_t nkeys = 0, care_sum = 0;
    int tmax = 0;  // largest threshold
};
struct KeyRef {
    const CompareKey *key;
    uint32_t index;  // written into the record; hits report it
};
int build_key_table(const std::vector<KeyRef> &keys, uint32_t P, KeyTable &t, std::string &err);

// Seed index of a key table built by build_key_table from `keys`: words =
// offs then ids (uint16 pairs packed in uint32).  False when the table does not qualify (IUPAC
// codes, two words, several care classes, thresholds too high for the
// comparable positions, or an index over the LDS budget).
struct SeedIndex {
    SeedDesc desc{};
    std::vector<uint32_t> words;
    uint32_t lds_bytes = 0;
};
bool build_seed_index(const std::vector<KeyRef> &keys, uint32_t P, const KeyTable &t, SeedIndex &s);

// Bit-set seed index (MaskDesc) of a key table built by build_key_table from
// `keys`, keeping records that match at least `minex` (1 or 2) ranges
// exactly.  False when the table does not qualify (IUPAC codes, two words,
// several care classes, more than 32 * kMaskMaxWords keys, thresholds too
// high for the comparable positions, or an index over the LDS budget).
struct MaskIndex {
    MaskDesc desc{};
    std::vector<uint32_t> words;
    uint32_t lds_bytes = 0;
};
bool build_mask_index(const std::vector<KeyRef> &keys, const KeyTable &t, int minex, MaskIndex &s);

// Bulge-family table of queries sharing padded pattern, D, R and PAM side.
struct BulgeTable {
    std::vector<uint32_t> words;
    uint32_t units = 0;  // care classes
    int nw = 1;
    bool gen = false;
    int ns = 0;                      // shifts per guide
    int32_t cb[kMaxShifts] = {};     // window position of bit 0 at each shift
    uint32_t gap[kMaxShifts] = {};   // RNA-bulge length of each shift's family
    bool rev = false;
    PatternDesc pat{};  // finder pattern of the group: the query with the least padding
    uint32_t dmin = 0;  // that query's DNA-bulge size (the padding every guide's own D extends)
    uint32_t dpad = 0;  // largest extra padding of a guide (largest D - dmin)
    uint64_t nguides = 0, nkeys = 0, care_sum = 0;
    int tmax = 0;  // largest threshold of an emitting variant
};
// Bit-set seed index of a bulge table (bulge_index_kernel): the guides'
// comparable positions split into `parts` ranges as in build_mask_index, a
// row per (range, 2-bit code) of the guides whose bases there equal the code
// (a position outside the guides' care inside a range matches every code).
// A variant of split j and RNA gap g compares guide positions below j at one
// window alignment and positions from j + g at another, so it breaks at most
// max(1, R) ranges; with threshold t at least parts - max(1, R) - t ranges
// match exactly at some alignment.  Rows are OR-ed over the alignments and a
// guide survives when at least `need` of them hold it, need per 32-guide word
// = parts - max(1, R) - the word's largest threshold (minex 3 layout of
// MaskDesc, no records: the bulge table's own records follow).  False for
// IUPAC guides, several care classes, two-word windows, more than 128 guides
// or thresholds too high for the comparable positions.
bool build_bulge_index(const BulgeTable &t, MaskIndex &s);

// Bit-sliced filter tables of a bulge table (bulge_slice_kernel, SliceDesc
// layout).  A variant of family u is within its threshold only if the
// guide's mismatches at alignment 0 and at alignment u, both counted over
// the care run, have at most (threshold + RNA gap) positions in common
// (cas-offinder.cpp:650-682: the left part of a split compares at one
// alignment, the right part at the other), and the no-bulge key only if
// its own count is at most its threshold, which bounds the first family's
// count too; so slot g passes family f when popc(M0 & Mf) <= T_f(g).  False
// for IUPAC guides, several care classes, a care run with holes or longer
// than 32, two-word windows, more than two bulge families (D + R > 2) or
// bulges above 2, more than
// kSliceMaxGuides guides or windows outside the staged tile.
struct SliceIndex {
    SliceDesc desc{};
    std::vector<uint32_t> words;
};
bool build_bulge_slices(const BulgeTable &t, const PatternDesc &pd, SliceIndex &s);

struct BulgeRef {
    const Query *q;
    uint32_t gbase;  // global index of the query's first key
};
// 1: built; 0: the queries do not fit the bulge-family kernel (use key tables).
// The pattern line of q without the bulge padding (D leading N for a
// forward PAM, D trailing N for a reversed one).
std::string unpadded_pattern(const Query &q);
// Bulge table of queries of one unpadded pattern and PAM side (their bulge
// sizes may differ; 1 + largest D + largest R <= kMaxShifts): 1 on success,
// 0 when the queries do not qualify (their keys then go to key tables).
int build_bulge_table(const std::vector<BulgeRef> &qs, BulgeTable &t);

}  // namespace cof
