// cli_common.h -- shared host code of the square_root / parallel_cc / approx_cut
// command lines: edge-list reader, CLICK generator, timing.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cac.h"

namespace cli {

// Reads the reference edge-list format (src/input/GraphInputIterator.cpp:3-36):
// the first line is ignored, then "n m", then m lines "from to weight".  The file
// is memory-mapped and parsed in parallel chunks.  Returns CAC_EIO with a message
// on a missing file or malformed content.  With CAC_EDGE_CACHE set (and not "0")
// in the environment, the parsed list is also written beside the file as
// <path>.cacbin (a 64-byte header and the cac_edge records), and later reads
// take it instead of parsing while the text file keeps its size and
// modification time; the text format stays the canonical input.
int read_edge_list(const std::string& path, uint32_t* n, std::vector<cac_edge>* edges, std::string* err);

// In-memory CLICK graph (src/input/CLICK.hpp:10-49): complete graph on n vertices,
// 10 clusters (i mod 10); weight = max(0, floor(N(8, 4))) inside a cluster and
// max(0, floor(N(4, 4))) across, the normal draws from Philox (Box-Muller) keyed
// by (seed, edge index) instead of mt19937 + std::normal_distribution.
std::vector<cac_edge> click_graph(uint32_t n, uint32_t seed);

double now_seconds();

// Parses like std::stof / std::stoi and reports a usage error on failure.
bool parse_float(const char* s, float* out);
bool parse_int(const char* s, int* out);

}  // namespace cli
