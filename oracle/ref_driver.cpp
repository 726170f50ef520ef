// ref_driver.cpp -- thin extern "C" entry points over the reference's own
// Karger-Stein library (src/karger-stein/, compiled from the reference tree by
// build_ref.sh into oracle/_ref/).  TEST INFRASTRUCTURE: used only by tests/
// and bench.py's cpu_baseline leg to pin the oracle against the reference code.
#include <cstdint>
#include <limits>

#include "co_mincut.h"

typedef unsigned long W;

static sparse_graph<W, edge<W>>* make_graph(stack_allocator* alloc, int n, long m, const uint64_t* e) {
  array<edge<W>> edges = alloc->allocate<edge<W>>(m);
  for (long i = 0; i < m; ++i) {
    edge<W> x;
    x.set_vertices((int)e[3 * i], (int)e[3 * i + 1]);
    x.weight = (W)e[3 * i + 2];
    edges[i] = x;
  }
  return new sparse_graph<W, edge<W>>(n, edges);
}

extern "C" {

int ref_ks_target(int s) { return comincut::target_number_of_vertices_for_contraction(s); }
double ref_ks_min_success(int s) { return comincut::min_success_in_one_trial(s); }
int ref_ks_trials(int n, double p) { return comincut::number_of_trials(n, p); }

// comincut::deterministic_minimum_cut (co_mincut.h:529-541): PR pass + Stoer-Wagner.
uint64_t ref_deterministic_mincut(int n, long m, const uint64_t* e) {
  stack_allocator alloc(0);
  sparse_graph<W, edge<W>>* g = make_graph(&alloc, n, m, e);
  W r = comincut::deterministic_minimum_cut(g);
  delete g;
  return r;
}

// comincut::minimum_cut(sparse_graph*, p, seed) (co_mincut.h:514-523): the
// Karger-Stein trials SequentialKargerSteinCut runs (dense switch at m >= n^2/32).
uint64_t ref_ks_mincut(int n, long m, const uint64_t* e, double p, int seed) {
  stack_allocator alloc(0);
  sparse_graph<W, edge<W>>* g = make_graph(&alloc, n, m, e);
  W r = comincut::minimum_cut(g, p, seed);
  delete g;
  return r;
}

}  // extern "C"
