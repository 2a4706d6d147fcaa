// Record formatting with the reference's output semantics
// (cas-offinder.cpp:319-328 headers, :394-455 records, :471-500 summary).
#pragma once

#include <map>
#include <string>

#include "devtypes.hpp"
#include "genome.hpp"
#include "query.hpp"

namespace cof {

using SummaryTable = std::map<std::string, unsigned>;

std::string header_lines();
// Appends the TSV records of hits[0..n) (all of one query) to `out`.
void format_hits(const HostGenome &g, const Query &q, const DevHit *hits, size_t n, std::string &out,
                 SummaryTable *summary);
std::string summary_text(const SummaryTable &s);

}  // namespace cof
