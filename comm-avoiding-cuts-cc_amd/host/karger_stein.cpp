// karger_stein -- drop-in for the reference's karger_stein command line
// (src/executables/karger_stein.cpp) on one MI355X.
//   karger_stein INPUT_FILE SEED
// Runs comincut::minimum_cut(graph, 0.90, seed) semantics through
// cac_karger_stein (dense engine when m >= n^2/32, else the sparse recursion)
// and prints input,seed,1,n,m,time,?,co-karger-stein,cut like the reference.
// Loops and zero-weight edges never cross a cut and are dropped before the call.
// time covers the upload and all trials, not the file parse.
#include <iostream>
#include <string>
#include <vector>

#include "cac.h"
#include "cli_common.h"

int main(int argc, char* argv[]) {
  if (argc != 3) {
    std::cout << "Usage: karger_stein INPUT_FILE SEED" << std::endl;
    return 1;
  }
  int seed_i = 0;
  if (!cli::parse_int(argv[2], &seed_i)) {
    std::cerr << "karger_stein: cannot parse SEED" << std::endl;
    return 1;
  }
  const uint32_t seed = (uint32_t)seed_i;
  uint32_t n = 0;
  std::vector<cac_edge> edges;
  std::string err;
  if (cli::read_edge_list(argv[1], &n, &edges, &err) != CAC_OK) {
    std::cerr << "karger_stein: " << err << std::endl;
    return 1;
  }
  const size_t m_file = edges.size();
  size_t k = 0;
  for (const cac_edge& e : edges)
    if (e.u != e.v && e.w != 0) edges[k++] = e;
  edges.resize(k);
  cac_ctx* ctx = nullptr;
  int rc = cac_ctx_create(0, &ctx);
  if (rc != CAC_OK) {
    std::cerr << "karger_stein: no GPU context (" << cac_strerror(rc) << ")" << std::endl;
    return 1;
  }
  int64_t cut = 0;
  uint32_t trials = 0;
  const double t0 = cli::now_seconds();
  if (edges.empty() || n < 2)
    cut = 0;  // no edge crosses any cut
  else
    rc = cac_karger_stein(ctx, edges.data(), edges.size(), n, 0.90, seed, &cut, &trials);
  const double t1 = cli::now_seconds();
  if (rc != CAC_OK) {
    std::cerr << "karger_stein: " << cac_strerror(rc) << ": " << cac_last_error(ctx) << std::endl;
    cac_ctx_destroy(ctx);
    return 1;
  }
  std::cout << argv[1] << "," << seed << "," << 1 << "," << n << "," << m_file << "," << (t1 - t0) << ","
            << "?"
            << ","
            << "co-karger-stein," << cut << std::endl;
  cac_ctx_destroy(ctx);
  return 0;
}
