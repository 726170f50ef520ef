// approx_cut -- drop-in for the reference's approx_cut command line
// (src/executables/approx_cut.cpp:5-44) on one MI355X.
//   approx_cut PROBABILITY INPUT_FILE|CLICK [SIZE] SEED
// Prints input,seed,p,n,m,cuttingTime,mpiTime,trials,approximation with p = 1.
#include <iostream>
#include <string>
#include <vector>

#include "cac.h"
#include "cli_common.h"

int main(int argc, char* argv[]) {
  if ((argc != 4) && (argc != 5)) {
    std::cout << "Usage: slimGraph-Light PROBABILITY INPUT_FILE|CLICK [SIZE] SEED" << std::endl;
    return 1;
  }
  float p = 0;
  int seed_i = 0;
  if (!cli::parse_float(argv[1], &p) || !cli::parse_int(argv[argc == 4 ? 3 : 4], &seed_i)) {
    std::cerr << "approx_cut: cannot parse PROBABILITY or SEED" << std::endl;
    return 1;
  }
  const uint32_t seed = (uint32_t)seed_i;
  uint32_t n = 0;
  std::vector<cac_edge> edges;
  if (std::string(argv[2]) == "CLICK") {
    if (argc != 5) {
      std::cerr << "approx_cut: CLICK needs SIZE" << std::endl;
      return 1;
    }
    n = (uint32_t)std::stoul(argv[3]);
    edges = cli::click_graph(n, seed);
  } else {
    std::string err;
    if (cli::read_edge_list(argv[2], &n, &edges, &err) != CAC_OK) {
      std::cerr << "approx_cut: " << err << std::endl;
      return 1;
    }
  }
  cac_ctx* ctx = nullptr;
  int rc = cac_ctx_create(0, &ctx);
  if (rc != CAC_OK) {
    std::cerr << "approx_cut: no GPU context (" << cac_strerror(rc) << ")" << std::endl;
    return 1;
  }
  uint64_t approx = 0;
  uint32_t trials = 0;
  const double t0 = cli::now_seconds();
  rc = cac_approx_cut(ctx, edges.data(), edges.size(), n, (double)p, seed, &approx, &trials);
  const double t1 = cli::now_seconds();
  if (rc != CAC_OK) {
    std::cerr << "approx_cut: " << cac_strerror(rc) << ": " << cac_last_error(ctx) << std::endl;
    cac_ctx_destroy(ctx);
    return 1;
  }
  std::cout << std::fixed;
  std::cout << argv[2] << "," << seed << "," << 1 << "," << n << "," << edges.size() << "," << (t1 - t0) << ","
            << 0.0 << "," << trials << "," << approx << std::endl;
  cac_ctx_destroy(ctx);
  return 0;
}
