// Resident genome: 2-bit bases + a non-ACGT bit plane, the host copy of what
// lives in HBM, plus what the record formatter needs to print the exact text.
//
// Global coordinate: the genome files are laid out one after another, each as
// the reference holds one file (records joined by ';', read_fasta.cpp:41 /
// read_twobit.cpp:123-127), with one ';' between files.  Position p of file f
// is text offset p - files[f].base of that file.  Every ';' (record and file
// separators, and a literal ';' inside a sequence line) is a "poison"
// position: a window containing one is not a candidate (cas-offinder.cl:64-67).
//
// Planes are little-endian bit arrays in 32-position words: bit (p & 31) of
// word p >> 5.  Base codes follow UCSC 2bit: T=0, C=1, A=2, G=3
// (read_twobit.cpp:30-42), so lo = C|G, hi = A|G and the complement is hi ^ 1.
// Non-ACGT positions have ot = 1 and lo = hi = 0; their byte lives in `exc`.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/casoff.h"

namespace cof {

constexpr int kPadWords = 16;  // zero words after the last used word

struct GenomeFile {
    std::string path;
    uint64_t base = 0;  // global position of text offset 0
    uint64_t len = 0;   // text length of this file
    std::vector<std::string> names;
    std::vector<uint64_t> chrpos;  // record start offsets in the file text
};

struct ExcRun {  // maximal run of one non-ACGT byte
    uint64_t start;
    uint64_t len;
    char ch;
};

class HostGenome {
public:
    uint64_t total = 0;
    std::vector<uint32_t> lo, hi, ot;
    std::vector<ExcRun> exc;
    std::vector<uint64_t> poison;
    std::vector<GenomeFile> files;

    void clear();
    uint64_t nwords() const { return (total + 31) >> 5; }
    char char_at(uint64_t pos) const;
    void text(uint64_t pos, uint64_t n, char *out) const;
    int file_of(uint64_t pos) const;  // -1 for a file separator
    uint64_t n_contigs() const;

    // Streaming builder (appends at `total`).
    void begin_file(const std::string &path);
    void end_file();
    void append(const char *s, size_t n);  // bytes already uppercased
    void append_parallel(const char *s, size_t n);  // same result, whole words packed by host threads
    void append_char(char c) { append(&c, 1); }
    void finish();  // pads the planes

private:
    uint32_t cur_lo_ = 0, cur_hi_ = 0, cur_ot_ = 0;
    void flush_word();
    void add_exc(uint64_t pos, char c);
};

// Readers: 0 loaded, 1 not this format (same contract as read_fasta.h:8 and
// read_twobit.h:9), -1 I/O error (err set).
int load_fasta(const std::string &path, HostGenome &g, std::string &err);
int load_twobit(const std::string &path, HostGenome &g, std::string &err);
// Directory (readdir order, DT_REG / DT_LNK entries, main.cpp:133-147) or one file.
int load_path(const std::string &path, HostGenome &g, int &nfiles, std::string &err, bool verbose);

// Synthetic genomes (seeded, deterministic).
void synth_genome(const cof_synth_spec &spec, HostGenome &g);
int synth_write(const cof_synth_spec &spec, const std::string &path, int format, std::string &err);

// Parallel-for over [0, n) on std::threads.
// fn over [0, n) split into at most n_host_threads() (and max_threads, when
// positive) ranges of at least `grain`, one thread each
void parallel_for(uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)> &fn,
                  int max_threads = 0);
int n_host_threads();

}  // namespace cof
