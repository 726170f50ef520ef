/*
 * oracle.h -- CPU restatement of the contraction core (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it.  The product path
 * (libcac.so) never links or calls it.  Every function states the reference
 * code it restates (paths relative to the reference repository root) and
 * what pins it (fixtures, closed forms, independent libraries, brute force).
 *
 * Integer arithmetic only; results are bit-exact specifications that the HIP
 * kernels must reproduce.
 */
#ifndef CAC_ORACLE_H
#define CAC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based PRNG (Salmon et al., SC'11) ----------------------------
 * Replaces the sequential sitmo::prng_engine stream (lib/prng_engine.hpp:195-212)
 * with a random-access generator.  Pinned by the published Random123 known-answer
 * vectors (4x32) and by numpy.random.Philox (4x64, same round structure). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* Uniform integer in [1, W] from a 64-bit word: 1 + floor(x * W / 2^64).
 * Replaces std::uniform_int_distribution<u64>(1, W)
 * (src/SequentialSquareRootCut.cpp:77-81, indexed_adjacency_matrix.hpp:117-118). */
uint64_t orc_uniform1(uint64_t x, uint64_t W);

/* ---- connected components -------------------------------------------------
 * Restates UnweightedIteratedSparseSampling::connectedComponents
 * (src/UnweightedIteratedSparseSampling.cpp:40-72) + prefixConnectedComponents
 * (src/IteratedSparseSampling.hpp:182-224): the partition is unique, so the
 * oracle is plain union-find over all edges.  mode 0: label = minimum vertex
 * id of the component; mode 1: label = dense rank of that minimum id (what the
 * reference's relabel loop at IteratedSparseSampling.hpp:210-220 yields).
 * Returns the number of components.  edges: m pairs (u, v). */
uint32_t orc_cc(uint32_t n, uint64_t m, const uint32_t* edges, int mode, uint32_t* labels);

/* R-MAT generator (spec of the synthetic input): draw d gives 'scale' quadrant
 * choices from philox4x32(ctr={d, level/4, 0, 0}, key={seed, 'RMAT'}), word
 * level%4, compared with cumulative thresholds ta < tab < tabc (u32 fixed
 * point).  Loops dropped, pairs canonicalised to (min, max), the first m
 * distinct pairs in draw order kept, output sorted by (u, v).
 * Returns the number of draws consumed. */
uint64_t orc_rmat(uint32_t scale, uint64_t m, uint32_t seed, uint32_t ta, uint32_t tab,
                  uint32_t tabc, uint32_t* edges_out);

/* ---- exact minimum cut ------------------------------------------------------
 * Stoer-Wagner (J. ACM 44(4), 1997) on a dense symmetric int64 matrix
 * (n x n row-major, zero diagonal).  Plays the role of the reference's
 * boost_stoer_wagner / utils/cut.py checkers and equals the value of the
 * reference base case deterministic_cut (co_mincut_base_case.hpp:52-144).
 * Pinned by brute force (n <= 20), networkx fixtures and hand-derived cuts. */
int64_t orc_stoer_wagner(uint32_t n, const int64_t* adj);
/* 2^(n-1) subset enumeration, n <= 24. */
int64_t orc_brute_mincut(uint32_t n, const int64_t* adj);

#ifdef __cplusplus
}
#endif
#endif
