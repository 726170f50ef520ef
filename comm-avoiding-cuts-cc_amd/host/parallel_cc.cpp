// parallel_cc -- drop-in for the reference's parallel_cc command line
// (src/executables/parallel_cc.cpp:6-66) on one MI355X.
//   parallel_cc INPUT_FILE SEED [ITERATIONS]
// One CSV line per iteration: input,seed,p,n,m,time,mpi,cc,#components with p = 1,
// mpi = 0; the seed increments per iteration as in the reference (the GPU
// algorithm is deterministic, so the count never depends on it).  time covers
// the device computation from the HBM-resident graph to labels, as the
// reference times only connectedComponents (parallel_cc.cpp:42-44).
//
// After loading, the edge list is turned into a symmetric CSR on the device
// once (cac_csr_build_dev, part of the load, like the reference's per-rank
// slices), and every iteration runs cac_cc_csr_dev on it.  Lists of 2^31 or
// more edges, or CAC_CC_INPUT=edges in the environment, run cac_cc_dev on the
// edge list instead.  Both give the same count and labels.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "cac.h"
#include "cli_common.h"

int main(int argc, char* argv[]) {
  if ((argc != 3) && (argc != 4)) {
    std::cout << "Usage: connected_components INPUT_FILE SEED [ITERATONS]" << std::endl;
    return 1;
  }
  int seed_i = 0, iterations = 1;
  if (!cli::parse_int(argv[2], &seed_i) || (argc == 4 && !cli::parse_int(argv[3], &iterations))) {
    std::cerr << "parallel_cc: cannot parse SEED or ITERATIONS" << std::endl;
    return 1;
  }
  uint32_t seed = (uint32_t)seed_i;
  uint32_t n = 0;
  std::vector<cac_edge> edges;
  std::string err;
  if (cli::read_edge_list(argv[1], &n, &edges, &err) != CAC_OK) {
    std::cerr << "parallel_cc: " << err << std::endl;
    return 1;
  }
  std::vector<cac_uedge> ue(edges.size());
  for (size_t i = 0; i < edges.size(); ++i) ue[i] = cac_uedge{edges[i].u, edges[i].v};
  for (const cac_uedge& x : ue)
    if (x.u >= n || x.v >= n) {
      std::cerr << "parallel_cc: edge endpoint >= n" << std::endl;
      return 1;
    }
  cac_ctx* ctx = nullptr;
  int rc = cac_ctx_create(0, &ctx);
  if (rc != CAC_OK) {
    std::cerr << "parallel_cc: no GPU context (" << cac_strerror(rc) << ")" << std::endl;
    return 1;
  }
  void *d_e = nullptr, *d_l = nullptr, *d_off = nullptr, *d_cols = nullptr;
  if (cac_dev_alloc(ctx, ue.size() * sizeof(cac_uedge), &d_e) || cac_dev_alloc(ctx, (size_t)n * 4, &d_l) ||
      cac_memcpy_h2d(ctx, d_e, ue.data(), ue.size() * sizeof(cac_uedge))) {
    std::cerr << "parallel_cc: " << cac_last_error(ctx) << std::endl;
    return 1;
  }
  const char* input = std::getenv("CAC_CC_INPUT");
  const bool use_csr = ue.size() < (1ull << 31) && !(input && std::strcmp(input, "edges") == 0);
  if (use_csr) {
    uint64_t nnz = 0;
    if (cac_dev_alloc(ctx, ((size_t)n + 1) * 4, &d_off) || cac_dev_alloc(ctx, 2 * ue.size() * 4 + 4, &d_cols) ||
        cac_csr_build_dev(ctx, (const cac_uedge*)d_e, ue.size(), n, (uint32_t*)d_off, (uint32_t*)d_cols, &nnz,
                          nullptr) ||
        cac_ctx_synchronize(ctx)) {
      std::cerr << "parallel_cc: " << cac_last_error(ctx) << std::endl;
      return 1;
    }
    cac_dev_free(ctx, d_e);  // the CSR holds every edge
    d_e = nullptr;
  }
  std::cout << std::fixed;
  for (int it = 0; it < iterations; ++it) {
    uint32_t ncomp = 0;
    const double t0 = cli::now_seconds();
    rc = use_csr ? cac_cc_csr_dev(ctx, (const uint32_t*)d_off, (const uint32_t*)d_cols, n, CAC_LABEL_RANK,
                                  (uint32_t*)d_l, &ncomp, nullptr)
                 : cac_cc_dev(ctx, (const cac_uedge*)d_e, ue.size(), n, CAC_LABEL_RANK, (uint32_t*)d_l, &ncomp,
                              nullptr);
    const double t1 = cli::now_seconds();
    if (rc != CAC_OK) {
      std::cerr << "parallel_cc: " << cac_strerror(rc) << ": " << cac_last_error(ctx) << std::endl;
      return 1;
    }
    std::cout << argv[1] << "," << seed << "," << 1 << "," << n << "," << edges.size() << "," << (t1 - t0) << ","
              << 0.0 << ","
              << "cc"
              << "," << ncomp << std::endl;
    seed++;
  }
  if (d_e) cac_dev_free(ctx, d_e);
  if (d_off) cac_dev_free(ctx, d_off);
  if (d_cols) cac_dev_free(ctx, d_cols);
  cac_dev_free(ctx, d_l);
  cac_ctx_destroy(ctx);
  return 0;
}
