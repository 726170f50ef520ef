// square_root -- drop-in for the reference's square_root command line
// (src/executables/square_root.cpp:5-51) on one MI355X.
//   square_root PROBABILITY INPUT_FILE|CLICK [SIZE] SEED
// Prints input,seed,p,n,m,cuttingTime,mpiTime,trials,high|low,cut with p = 1
// (one GPU), mpiTime = 0 and the low-concurrency variant (SquareRootCut.cpp:11-13
// always chooses it when p < 2 T).  cuttingTime covers upload and all trials,
// not the file parse, like the reference's timer (SquareRootCut.cpp:114-141).
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "cac.h"
#include "cli_common.h"

int main(int argc, char* argv[]) {
  if ((argc != 4) && (argc != 5)) {
    std::cout << "Usage: square_root PROBABILITY INPUT_FILE|CLICK [SIZE] SEED" << std::endl;
    return 1;
  }
  float p = 0;
  int seed_i = 0;
  if (!cli::parse_float(argv[1], &p) || !cli::parse_int(argv[argc == 4 ? 3 : 4], &seed_i)) {
    std::cerr << "square_root: cannot parse PROBABILITY or SEED" << std::endl;
    return 1;
  }
  const uint32_t seed = (uint32_t)seed_i;
  uint32_t n = 0;
  std::vector<cac_edge> edges;
  if (std::string(argv[2]) == "CLICK") {
    if (argc != 5) {
      std::cerr << "square_root: CLICK needs SIZE" << std::endl;
      return 1;
    }
    n = (uint32_t)std::stoul(argv[3]);
    edges = cli::click_graph(n, seed);
  } else {
    std::string err;
    if (cli::read_edge_list(argv[2], &n, &edges, &err) != CAC_OK) {
      std::cerr << "square_root: " << err << std::endl;
      return 1;
    }
  }
  cac_ctx* ctx = nullptr;
  int rc = cac_ctx_create(0, &ctx);
  if (rc != CAC_OK) {
    std::cerr << "square_root: no GPU context (" << cac_strerror(rc) << ")" << std::endl;
    return 1;
  }
  int64_t cut = 0;
  uint32_t trials = 0;
  const double t0 = cli::now_seconds();
  rc = cac_mincut(ctx, edges.data(), edges.size(), n, (double)p, seed, &cut, &trials);
  const double t1 = cli::now_seconds();
  if (rc != CAC_OK) {
    std::cerr << "square_root: " << cac_strerror(rc) << ": " << cac_last_error(ctx) << std::endl;
    cac_ctx_destroy(ctx);
    return 1;
  }
  std::cout << std::fixed;
  std::cout << argv[2] << "," << seed << "," << 1 << "," << n << "," << edges.size() << "," << (t1 - t0) << ","
            << 0.0 << "," << trials << ","
            << "low"
            << "," << cut << std::endl;
  cac_ctx_destroy(ctx);
  return 0;
}
