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
/* The same labels and count on `threads` threads (0 = every online core): the
 * multithreaded CPU baseline of the connected-components benchmark. */
uint32_t orc_cc_mt(uint32_t n, uint64_t m, const uint32_t* edges, int mode, uint32_t* labels, int threads);

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



/* ---- trial-count and size rules (floating point exactly as the reference) ---
 * comincut::target_number_of_vertices_for_contraction (src/karger-stein/co_mincut.cpp:18-20),
 * min_success_in_one_trial (:32-40), number_of_trials (:43-74),
 * SquareRootCut::intermediate_size (src/SquareRootCut.cpp:63-70), numberOfTrials (:76-83),
 * ApproximateCut::numberOfTrials (src/ApproximateCut.cpp:82-87). */
uint32_t orc_ks_target(uint32_t s);
double orc_ks_min_success(uint32_t s);
uint32_t orc_ks_trials(uint32_t n, double p);
uint32_t orc_intermediate_size(uint32_t n, uint64_t m);
uint32_t orc_sqrt_trials(uint32_t n, uint64_t m, double p);
uint32_t orc_approx_trials(double p);

/* ---- dense Karger-Stein (one trial) -----------------------------------------
 * Restates comincut::minimum_cut_trials_start(dense) + recursive_contraction_inner
 * (co_mincut.h:370-390, 176-223), contract_graph (co_mincut.h:38-66),
 * indexed_adjacency_matrix::{contract_random_edge, contract_edge, pr_pass, compact}
 * (indexed_adjacency_matrix.hpp:80-98, 137-191) and lazy_adjacency_matrix::
 * {contract_edge, compact} (lazy_adjacency_matrix.hpp:120-216), with:
 *   - picks keyed by Philox(ctr={trial, node, pick, 0}, key={seed,'KARG'}): row i
 *     = first index with prefix(capacity) >= 1 + floor(lo64*W/2^64), column k =
 *     first index with prefix(row i) >= 1 + floor(hi64*cap_i/2^64), j = rep[k];
 *   - natural-order prefix instead of the sum_tree order;
 *   - every node of the trial uses the trial's initial bound (min weighted degree);
 *   - leaves (<= 128 vertices) return min(bound, exact min cut).
 * Returns the raw cut of the trial: min(bound, every leaf).  adj is n x n,
 * symmetric, zero diagonal, non-negative. */
int64_t orc_ks_trial(uint32_t n, const int64_t* adj, uint32_t seed, uint32_t trial);
int64_t orc_min_degree(uint32_t n, const int64_t* adj);
/* The first pick of node nid (max, min order), for the selection-distribution test. */
void orc_ks_first_pick(uint32_t n, const int64_t* adj, uint32_t seed, uint32_t trial, uint32_t nid, uint32_t* a,
                       uint32_t* b);

/* ---- weighted iterated sampling (one sqrt-cut trial's shrink phase) ----------
 * Restates SequentialSquareRootCut::iteratedSampling (src/SequentialSquareRootCut.cpp:21-67)
 * with AdjacencyListGraph::{weaklyContractEdge, finalizeContractionPhase}
 * (src/AdjacencyListGraph.cpp:54-69, 87-112): rounds of |V| draws (with
 * replacement) proportional to weight over the edges whose endpoints are still
 * in different super-vertices, draw d of round r keyed Philox(ctr={trial, r, d, 0},
 * key={seed,'SAMP'}) lo64; contraction in draw order, stopping as soon as |V| == t.
 * labels[v] = smallest original id of v's super-vertex.  Returns |V| at the end
 * (== t unless the graph has more than t components). edges: m x (u, v, w). */
uint32_t orc_sample_contract(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed,
                             uint32_t trial, uint32_t* labels);
/* The same with draws_per_round draws in every round (0: |V|, the rule above).
 * The high-concurrency rule (WeightedIteratedSparseSampling::shrink,
 * src/WeightedIteratedSparseSampling.cpp:4-24, 236-279) draws orc_hc_draws(n)
 * = (unsigned) powf(n, 1.05f) per round (.hpp:66-84). */
uint32_t orc_sample_contract_rule(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed,
                                  uint32_t trial, uint32_t draws_per_round, uint32_t* labels);
uint32_t orc_hc_draws(uint32_t n0);
/* Contracted multigraph -> dense matrix over the ranks of the canonical labels
 * (AdjacencyListGraph::compact, src/AdjacencyListGraph.cpp:34-52, renamed by rank
 * of the minimum id instead of first appearance).  out: k x k, k = distinct labels. */
uint32_t orc_densify(uint32_t n, uint64_t m, const uint64_t* edges, const uint32_t* labels, int64_t* out);
/* One low-concurrency sqrt-cut trial: sample_contract to t, densify, one dense KS
 * trial with the same seed and trial id (SequentialSquareRootCut::compute,
 * src/SequentialSquareRootCut.cpp:4-19).  For t > 8192 (beyond the dense
 * engine) the contracted graph goes, as an edge list ranked by canonical label,
 * loops dropped, parallel edges summed and sorted by (u, v), through one sparse
 * Karger-Stein trial (orc_ks_sparse_trial, node 1); a disconnected graph gives 0. */
int64_t orc_sqrt_trial(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed, uint32_t trial);
int64_t orc_sqrt_trial_rule(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed, uint32_t trial,
                            uint32_t draws_per_round);

/* ---- approximate minimum cut ------------------------------------------------
 * Restates ApproximateCut::run (src/ApproximateCut.cpp:10-79): for i = 1 ..
 * ceil(ln W) (W = total weight), copies j < T (T = orc_approx_trials(p)) keep edge e
 * iff binomial(w_e, 2^-i) > 0, here exactly: x >= Q_i(w_e), x = lo64 of
 * Philox(ctr={i, j, e_lo, e_hi}, key={seed,'APRX'}), Q_i(w) = (1 - 2^-i)^w in
 * 0.64 fixed point by square-and-multiply with floor(a*b/2^64); stop at the first
 * i whose union of copies has more than T components.  *approx = 2^i (1 if none),
 * *trials = T * i. */
uint64_t orc_keep_threshold(uint32_t i, uint64_t w);
void orc_approx_cut(uint32_t n, uint64_t m, const uint64_t* edges, double p, uint32_t seed, uint64_t* approx,
                    uint32_t* trials);

/* ---- sparse Karger-Stein ---------------------------------------------------
 * Restates the sparse recursion of comincut (co_mincut.h:74-155 contraction,
 * :176-223 recursion, :280-293 sparse dispatch) with these rules of this build:
 *   - the exponential race uses integer keys orc_race_key(x, w) from
 *     Philox(ctr={trial, nid, e_lo, e_hi}, key={seed,'SPRS'}) word pair 0-1,
 *     edges taken in (key, index) order until t' components remain;
 *   - the child keeps the edges in their order, relabelled by component rank,
 *     loops dropped, parallel edges not merged;
 *   - a node with n <= 2048, or with m >= n^2/32 and n <= 8192, is densified and continues as the
 *     dense recursion (orc_ks_trial's nodes) under the same node id;
 *   - every node uses the trial's bound, the minimum weighted degree.
 * orc_ks_sparse_trial returns min(bound, every leaf) of one trial.
 * edges: m x (u, v, w), u < v. */
uint64_t orc_race_key(uint64_t x, uint64_t w);
uint64_t* orc_sparse_contract(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t tgt, uint32_t seed,
                              uint32_t trial, uint32_t nid, uint64_t* m_out);
int64_t orc_ks_sparse_trial(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t seed, uint32_t trial);

#ifdef __cplusplus
}
#endif
#endif /* CAC_ORACLE_H */
