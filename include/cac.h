/*
 * cac.h -- C ABI of the MI355X (gfx950) contraction core: connected components,
 * weighted sampling contraction, dense Karger-Stein and the min-cut drivers.
 *
 * Conventions
 *   - Every entry point returns CAC_OK (0) or a negative CAC_E* code; no C++
 *     exception crosses this boundary.  cac_strerror() names a code and
 *     cac_last_error() gives the message of the last failure on a context.
 *   - Functions without the _dev suffix take and return HOST memory; the caller
 *     owns every buffer, the context owns device scratch.  _dev variants take
 *     device pointers and an optional hipStream_t passed as void* (NULL = the
 *     context's own stream) and return once the results are on the host side
 *     of the call (scalars) / enqueued (device arrays) as documented per call.
 *   - A context is bound to one device and is not thread-safe: use one per
 *     host thread.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef CAC_H
#define CAC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define CAC_OK 0
#define CAC_EINVAL (-1)   /* bad argument: null pointer, vertex id >= n, size out of range */
#define CAC_ENOMEM (-2)   /* device or host allocation failed */
#define CAC_EDEVICE (-3)  /* HIP runtime error */
#define CAC_ENOTCONN (-4) /* more components than a contraction target (a min cut of a disconnected graph is 0) */
#define CAC_ERANGE (-5)   /* a weight sum would overflow int64 */
#define CAC_EIO (-6)      /* edge-list file cannot be opened or parsed */

/* ---- edge records --------------------------------------------------------
 * cac_uedge mirrors UnweightedGraph::Edge {unsigned from, to} (src/UnweightedGraph.hpp);
 * cac_edge mirrors AdjacencyListGraph::Edge {unsigned from, to; unsigned long weight}
 * (src/AdjacencyListGraph.hpp:13-50), 16 bytes. */
typedef struct cac_uedge {
  uint32_t u, v;
} cac_uedge;

typedef struct cac_edge {
  uint32_t u, v;
  uint64_t w;
} cac_edge;

typedef struct cac_ctx cac_ctx;

/* ---- context -------------------------------------------------------------
 * Replaces the per-rank MPI setup of the reference CLIs (src/executables/,
 * MPI_Init / communicator); one context per GPU and host thread. */
int cac_ctx_create(int device, cac_ctx** out);
int cac_ctx_destroy(cac_ctx* ctx);
const char* cac_strerror(int code);
const char* cac_last_error(const cac_ctx* ctx);
/* The context's HIP stream (hipStream_t as void*). */
void* cac_ctx_stream(cac_ctx* ctx);
int cac_ctx_synchronize(cac_ctx* ctx);

/* Device buffers for the _dev entry points (plain hipMalloc/hipMemcpy on the
 * context's device; no counterpart in the reference, which has no device). */
int cac_dev_alloc(cac_ctx* ctx, size_t bytes, void** dptr);
int cac_dev_free(cac_ctx* ctx, void* dptr);
int cac_memcpy_h2d(cac_ctx* ctx, void* dst, const void* src, size_t bytes);
int cac_memcpy_d2h(cac_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a streaming kernel (16-byte aligned pointers),
 * enqueued on `stream` (NULL = the context's stream) without synchronising.
 * Reads and writes exactly `bytes` bytes, so it also calibrates HBM byte
 * counters.  No counterpart in the reference. */
int cac_memcpy_d2d(cac_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);

/* Launch shapes picked by the autotuner (comm-avoiding-cuts-cc_amd/autotune.py),
 * held per context; results never depend on them.  No counterpart in the
 * reference, whose launch shape is one MPI rank per core.
 * CAC_TUNE_KS_CONTRACT_BLOCK: threads per workgroup of the dense Karger-Stein
 * contraction for children of s vertices, keyed by row-size class: key 512
 * holds s <= 512, keys 1024, 2048 and 4096 hold s in (key/2, key] (s above 4096
 * always uses 1024).  value 0 restores the built-in rule (the smallest power
 * of two >= s/8, at least 64); otherwise a power of two from max(64, key/8) to
 * 1024, CAC_EINVAL else. */
#define CAC_TUNE_KS_CONTRACT_BLOCK 1
/* CAC_TUNE_SAMPLE_OCCUPANCY: 1024-thread sampling workgroups per CU for a
 * launch of more square-root trials than CUs, keyed by the class of the
 * launch's largest edge count m: key 16 holds m < 2^16, 20 m < 2^20, 24
 * m < 2^24, 32 the rest.  value 0 restores the built-in 2 (64 registers per
 * thread, some spilled); 1 gives the workgroup the registers it wants;
 * CAC_EINVAL else. */
#define CAC_TUNE_SAMPLE_OCCUPANCY 2
/* CAC_TUNE_WIDE_SAMPLING: key 0; value v = log2 of the smallest edge count
 * whose square-root trials run in the wide form (each trial's sampling rounds
 * and densification as grid-wide kernels instead of one workgroup per trial).
 * v = 0 restores the built-in 20 (2^20 edges), v = 63 never uses the wide form,
 * 1 <= v <= 62 sets 2^v; CAC_EINVAL else.  Cuts do not depend on it. */
#define CAC_TUNE_WIDE_SAMPLING 3
/* CAC_TUNE_KS_COMPACT_BLOCK: threads per workgroup of the dense Karger-Stein
 * compaction for children of t' vertices, keyed by class: key 512 holds
 * t' <= 512, keys 1024, 2048, 4096 and 8192 hold t' in (key/2, key].  value 0
 * restores the built-in 512; otherwise 256, 512 or 1024, CAC_EINVAL else.
 * Results never depend on it. */
#define CAC_TUNE_KS_COMPACT_BLOCK 4
int cac_tune_set(cac_ctx* ctx, int knob, uint32_t key, int32_t value);
int cac_tune_get(cac_ctx* ctx, int knob, uint32_t key, int32_t* value);

/* Per-kernel timing with HIP events on the stream each kernel is launched on.
 * No counterpart in the reference beyond TimeUtils::measure (src/utils.hpp:21-40). */
int cac_prof_enable(cac_ctx* ctx, int on);
int cac_prof_reset(cac_ctx* ctx);
int cac_prof_count(cac_ctx* ctx);
/* Synchronizes and returns name and duration (ms) of recorded launch i. */
int cac_prof_get(cac_ctx* ctx, int i, const char** name, float* ms);

/* ---- connected components ------------------------------------------------
 * Replaces UnweightedIteratedSparseSampling::connectedComponents(std::vector<unsigned>&)
 * (src/UnweightedIteratedSparseSampling.hpp:19, .cpp:40-72).  Undirected edges;
 * loops and duplicates allowed; every id must be < n (CAC_EINVAL otherwise).
 * label_mode CAC_LABEL_MIN_ID: labels[v] = smallest vertex id of v's component.
 * label_mode CAC_LABEL_RANK:   labels[v] = rank of that smallest id among all
 *                              components (the reference's [0, count) labels).
 * *ncomp_out = number of components (the value parallel_cc prints). */
#define CAC_LABEL_MIN_ID 0
#define CAC_LABEL_RANK 1
int cac_cc(cac_ctx* ctx, const cac_uedge* edges, uint64_t m, uint32_t n, int label_mode,
           uint32_t* labels_out, uint32_t* ncomp_out);
/* Device variant: d_labels (n entries) is also the working array.  Ids are not
 * validated (the caller guarantees < n).  Blocks until *ncomp_out is known. */
int cac_cc_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode,
               uint32_t* d_labels, uint32_t* ncomp_out, void* stream);
/* cac_cc_dev without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, labels and count as cac_cc_dev; its scratch (counters, the
 * giant-component bitmap, the rank prefix) is a static device array of the
 * library, about 6 MiB, which every GPU the library runs on holds until the
 * process ends.  n <= CAC_CC_STATIC_MAX_N (CAC_EINVAL beyond).  All calls on
 * one device share that array: calls that overlap on one GPU (other streams,
 * other threads or contexts) race and must be serialised by the caller. */
#define CAC_CC_STATIC_MAX_N (1u << 25)
int cac_cc_static_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode,
                      uint32_t* d_labels, uint32_t* ncomp_out, void* stream);

/* ---- connected components on a symmetric CSR -------------------------------
 * The same labels as cac_cc_dev (CAC_LABEL_MIN_ID / CAC_LABEL_RANK, component
 * count), for a graph held as a symmetric CSR: d_offsets[n + 1], d_cols[nnz],
 * every undirected edge in both adjacency lists (loops and repeats allowed).
 * Afforest (Sutton, Ben-Nun and Barak, IPDPS 2018): the low-id vertices link
 * with their first two neighbours and every other vertex with its first, the
 * root of the largest component is guessed by sampling, and only vertices
 * outside it link their other neighbours, so most of the adjacency of a graph
 * with a giant component is never read.  Replaces the same
 * connectedComponents call as cac_cc_dev (src/UnweightedIteratedSparseSampling.cpp:40-72)
 * for callers that hold a CSR.  Ids are not validated. */
int cac_cc_csr_dev(cac_ctx* ctx, const uint32_t* d_offsets, const uint32_t* d_cols, uint32_t n, int label_mode,
                   uint32_t* d_labels, uint32_t* ncomp_out, void* stream);

/* cac_cc_csr_dev through a captured HIP graph: the first call with a given
 * (d_offsets, d_cols, n, label_mode, d_labels, stream) records the kernel
 * sequence of cac_cc_csr_dev into a graph held by the context; that call and
 * every later one with the same arguments replay it as one launch (a call with
 * other arguments, or after the context's scratch moved, captures again).  The
 * graph holds kernel nodes only.  Same labels and count as cac_cc_csr_dev.  For
 * callers that repeat the step on resident data (the serving pattern); no
 * counterpart in the reference. */
int cac_cc_csr_graph_dev(cac_ctx* ctx, const uint32_t* d_offsets, const uint32_t* d_cols, uint32_t n,
                         int label_mode, uint32_t* d_labels, uint32_t* ncomp_out, void* stream);
/* Symmetric CSR of an undirected device edge list (ids < n, checked:
 * CAC_EINVAL otherwise; fewer than 2^31 edges): both orientations of every non-loop edge, each adjacency list sorted
 * by neighbour id; d_offsets[n + 1], d_cols[2 m] (the first *nnz_out used).
 * Set-up work (radix sort), no counterpart in the reference. */
int cac_csr_build_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, uint32_t* d_offsets,
                      uint32_t* d_cols, uint64_t* nnz_out, void* stream);

/* ---- dense Karger-Stein --------------------------------------------------
 * Replaces comincut::minimum_cut(adjacency_matrix<T>*, double, int) and
 * minimum_cut_try (src/karger-stein/co_mincut.h:466-490; trials co_mincut.h:315-342).
 * adj: n x n row-major int64, symmetric, zero diagonal, non-negative, n >= 2,
 * n <= 8192.  Runs trials first_trial .. first_trial+ntrials-1; trial t's draws are
 * Philox-4x32-10(ctr = {t, node, pick, 0}, key = {seed, 'KARG'}).  Every trial's
 * upper bound is the minimum weighted degree (co_mincut.h:381).
 * raw_cuts[i] (optional, ntrials entries) = min(bound, every leaf of trial i);
 * *min_cut = min over the trials.  The number of trials for success probability
 * p is cac_ks_trials(n, p) (comincut::number_of_trials, co_mincut.cpp:43-74). */
int cac_ks_dense(cac_ctx* ctx, const int64_t* adj, uint32_t n, uint32_t seed, uint32_t first_trial,
                 uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut);
/* Device variant: d_adj in device memory (not validated); results on the host. */
int cac_ks_dense_dev(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, uint32_t seed, uint32_t first_trial,
                     uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut, void* stream);

/* ---- sparse Karger-Stein ------------------------------------------------------
 * Replaces comincut::minimum_cut(sparse_graph<T>*, double, int) on a graph that
 * stays sparse: contract_graph / sparse_contract_edges (src/karger-stein/
 * co_mincut.h:74-155) inside recursive_contraction (:176-223, :280-293).
 * edges: u != v, ids < n, 1 <= w < 2^40, total weight < 2^61, either order,
 * parallel edges allowed; fewer than 2^32 - 1 edges.  A node of n vertices is contracted to
 * t' = max(ceil(n / sqrt 2) + 1, 128) vertices by uniting its edges in
 * (race key, index) order, the race key of edge e of the node with id nid in
 * trial t being cac_race_key(x, w), x = the low 64 bits of
 * Philox-4x32-10(ctr = {t, nid, e_lo, e_hi}, key = {seed, 'SPRS'}); the children
 * of node nid are 2 nid and 2 nid + 1 (the root is 1).  A child keeps the
 * node's edges in order, relabelled by the rank of each component's minimum
 * vertex id, loops dropped, parallel edges not merged.  A node with n <= 2048,
 * or with m >= n^2 / 32 and n <= 8192, continues in the dense engine
 * (cac_ks_dense rules) under its own node id.  The bound of every trial is the
 * minimum weighted degree; a disconnected graph has cut 0.  raw_cuts (optional,
 * ntrials entries) and *min_cut as for cac_ks_dense. */
int cac_ks_sparse(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t seed, uint32_t first_trial,
                  uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut);
/* One sparse contraction, the step above for node nid of trial `trial`: the
 * child edge list of (edges, m, n) contracted to tgt vertices (tgt >= n: no
 * contraction, loops none).  Edges must have u < v < n and 1 <= w < 2^40.
 * *m_out = child edges; CAC_ERANGE if out_cap is smaller, CAC_ENOTCONN if all
 * edges leave more than tgt components.  Replaces sparse contract_graph
 * (co_mincut.h:131-155) without its PR pass. */
int cac_sparse_contract(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t tgt, uint32_t seed,
                        uint32_t trial, uint32_t nid, cac_edge* out, uint64_t out_cap, uint64_t* m_out);
/* Replaces SequentialKargerSteinCut::compute (src/SequentialKargerSteinCut.cpp:6-27)
 * with comincut::minimum_cut_trials_start's dense switch (co_mincut.h:396-414):
 * T = cac_ks_trials(n, p) trials (ids 0..T-1; p = 0 gives one trial); the dense
 * engine when m >= n^2 / 32 and n <= 8192, else cac_ks_sparse (which hands a
 * node over to the dense engine once it is small and dense enough).
 * *cut = minimum, *trials = T. */
int cac_karger_stein(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                     int64_t* cut, uint32_t* trials);
/* The race key of the sparse contraction: floor(2^20 * L / w) with
 * L = -log2(((x >> 11) + 1) / 2^53) in Q32.32 fixed point, the fraction of the
 * logarithm by 32 steps of mantissa squaring.  Exposed for tests. */
uint64_t cac_race_key(uint64_t x, uint64_t w);

/* ---- sizes and trial counts ------------------------------------------------
 * Same floating point as the reference: comincut::number_of_trials
 * (src/karger-stein/co_mincut.cpp:43-74), SquareRootCut::intermediate_size and
 * numberOfTrials (src/SquareRootCut.cpp:63-83, base_case_multiplier 2).  The
 * reference parses p with std::stof (src/executables/square_root.cpp:12); pass
 * the float value to reproduce its trial counts. */
uint32_t cac_ks_trials(uint32_t n, double p);
uint32_t cac_intermediate_size(uint32_t n, uint64_t m);
uint32_t cac_sqrt_trials(uint32_t n, uint64_t m, double p);

/* ---- weighted iterated sampling ---------------------------------------------
 * Replaces SequentialSquareRootCut::iteratedSampling + AdjacencyListGraph::
 * finalizeContractionPhase / compact (src/SequentialSquareRootCut.cpp:21-67,
 * src/AdjacencyListGraph.cpp:34-112) for one trial: rounds of |V| draws
 * proportional to weight among the edges joining different super-vertices, draw d
 * of round r from Philox-4x32-10(ctr = {trial, r, d, 0}, key = {seed, 'SAMP'}),
 * r = 1 + floor(x * W / 2^64) on the 64-bit word x, edge = first inclusive weight
 * prefix >= r; contraction in draw order stops as soon as |V| == t.
 * vmap_out[v] (n entries) = id in [0, *n_out) of v's super-vertex, ordered by the
 * smallest original vertex of each super-vertex; merged_out receives the merged
 * edge list (u < v, sorted by (u, v), parallel weights summed; *m_out entries;
 * CAC_ERANGE if merged_cap is smaller).  *n_out == t unless the graph has more
 * than t components. */
int cac_sample_contract(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t trial, uint32_t* vmap_out, uint32_t* n_out, cac_edge* merged_out,
                        uint64_t merged_cap, uint64_t* m_out);

/* Sampling rule of the shrink phase for every sampling call on this context
 * (cac_sample_contract, cac_sqrt_cut_trials, cac_mincut, cac_mincut_batch):
 *   CAC_SAMPLE_LC (the default): |V| draws per round, the low-concurrency rule
 *     of SequentialSquareRootCut::iteratedSampling (src/SequentialSquareRootCut.cpp:37-48);
 *   CAC_SAMPLE_HC: cac_hc_draws(n) draws per round, the high-concurrency rule of
 *     WeightedIteratedSparseSampling::shrink / edgesToSamplePerProcessor
 *     (src/WeightedIteratedSparseSampling.cpp:4-24, 236-279, .hpp:66-84).
 * The reference takes the high-concurrency path when its ranks number at least
 * twice the trials (SquareRootCut::lowConcurrency, src/SquareRootCut.cpp:11-13).
 * Draw d of round r is the same Philox word under both rules; CAC_EINVAL for
 * another value. */
#define CAC_SAMPLE_LC 0
#define CAC_SAMPLE_HC 1
int cac_set_sampling_rule(cac_ctx* ctx, int rule);
/* (unsigned) powf(n0, 1.05f): the reference's draws per round for n0 vertices
 * (src/WeightedIteratedSparseSampling.hpp:66-68, epsilon = 0.1f). */
uint32_t cac_hc_draws(uint32_t n0);

/* ---- square-root min cut ----------------------------------------------------
 * One low-concurrency trial = SequentialSquareRootCut(graph, engine, t).compute()
 * (src/SequentialSquareRootCut.cpp:4-19): the sampling contraction above, then
 * one dense Karger-Stein trial (same seed, same trial id) on the merged graph.
 * cac_sqrt_cut_trials returns the raw cut of trials first_trial.. (ntrials entries). */
int cac_sqrt_cut_trials(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t first_trial, uint32_t ntrials, int64_t* raw_cuts);
/* Replaces SquareRootCut::runMaster / runLowConcurrencyMaster (src/SquareRootCut.cpp:15-25,
 * 86-148): t = cac_intermediate_size(n, m), T = cac_sqrt_trials(n, m, p) trials
 * (ids 0..T-1) on one GPU, *cut = minimum raw cut, *trials = T.  Edges must have
 * ids < n and weights whose sum stays below 2^61. */
int cac_mincut(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed, int64_t* cut,
               uint32_t* trials);
int cac_mincut_dev(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                   int64_t* cut, uint32_t* trials, void* stream);

/* ---- batched square-root min cut ---------------------------------------------
 * Many independent graphs in one call.  The reference solves one graph per MPI
 * job (src/SquareRootCut.cpp:15-25, src/executables/square_root.cpp); there is
 * no batch interface to replace.  Graph g's edges are
 * d_edges[edge_offset .. edge_offset + m) with ids local to the graph (< n).
 * Graph g runs square-root-cut trials 0 .. T_g - 1 exactly as cac_mincut would
 * on that graph alone (t = cac_intermediate_size(n, m)), so every cut equals the
 * one-graph result with the same T_g:
 *   trials > 0:  T_g = trials;   trials == 0:  T_g = cac_sqrt_trials(n, m, p).
 * cuts[g] = minimum raw cut over the graph's trials; trials_out[g] (optional) =
 * T_g.  The (graph, trial) jobs of all graphs share launches, in chunks sized to
 * the free device memory. */
typedef struct cac_graph_desc {
  uint64_t edge_offset;
  uint64_t m;
  uint32_t n;
  uint32_t trials;
} cac_graph_desc;
int cac_mincut_batch(cac_ctx* ctx, const cac_edge* edges, const cac_graph_desc* descs, uint32_t count, double p,
                     uint32_t seed, int64_t* cuts, uint32_t* trials_out);
/* Device variant: d_edges in device memory (ids not validated); descs, cuts and
 * trials_out on the host; returns when the cuts are on the host. */
int cac_mincut_batch_dev(cac_ctx* ctx, const cac_edge* d_edges, const cac_graph_desc* descs, uint32_t count, double p,
                         uint32_t seed, int64_t* cuts, uint32_t* trials_out, void* stream);

/* ---- serving loop ------------------------------------------------------------
 * A query server on one GPU (no counterpart in the reference, whose executables
 * solve one graph per MPI job).  W worker threads, each with its own context
 * and stream, admit queries first come first served when their device-memory
 * footprint fits beside the work in flight (hbm_budget bytes in total), and
 * run up to max_batch pending queries of one kind as one launch group: CC
 * graphs side by side in one CC pipeline, the min-cut trials of all queries
 * (each with its own seed) in shared sampling / Karger-Stein launches.  Groups
 * on different workers overlap on the device.  Results do not depend on
 * batching:
 *   CC:      ncomp and labels (canonical minimum ids) equal cac_cc's;
 *   min cut: cut = min over trials 0 .. T-1 of the square-root cut with
 *            (t, T) from cac_serve_mincut_plan, i.e. the minimum of
 *            cac_sqrt_cut_trials(edges, n, t, seed, 0, T); with trials = 0
 *            (the reference's rule) that is cac_mincut(edges, n, p, seed).
 * A query's edges (and labels_out) are host memory that must stay valid until
 * its result is polled.  Invalid queries (ids >= n, weights >= 2^40, weight sum
 * >= 2^61, footprint > budget) come back with a non-zero status. */
#define CAC_QUERY_CC 0
#define CAC_QUERY_MINCUT 1
typedef struct cac_query {
  int32_t kind;            /* CAC_QUERY_CC or CAC_QUERY_MINCUT */
  uint32_t n;
  uint64_t m;
  const cac_edge* edges;   /* m records; CC ignores w */
  double p;                /* min cut: success probability */
  uint32_t seed;           /* min cut: Philox seed */
  uint32_t trials;         /* min cut: 0 = the reference's rule for p, k > 0 = k trials (cac_serve_mincut_plan) */
  uint32_t* labels_out;    /* CC: optional, n labels (minimum id of the component) */
  uint64_t graph;          /* 0: edges/n/m above; else a cac_server_register_graph handle (edges, n, m ignored) */
} cac_query;
typedef struct cac_query_result {
  uint64_t id;             /* as returned by cac_server_submit */
  int32_t status;          /* CAC_OK or a CAC_E* code */
  uint32_t ncomp;          /* CC: number of components */
  int64_t cut;             /* min cut */
  uint32_t t, trials;      /* min cut: intermediate size and trials run */
  double submit_s, admit_s, done_s;  /* server clock (seconds since create) */
  uint32_t batch_size;     /* queries in the launch group it ran in */
  uint32_t reserved;
} cac_query_result;
typedef struct cac_server cac_server;
/* Served min cut: the intermediate size t and trial count T of a query.
 * trials_req == 0: the reference's rule for success probability p, as cac_mincut
 *   runs it: t = cac_intermediate_size(n, m), T = cac_sqrt_trials(n, m, p)
 *   (SquareRootCut::intermediate_size / numberOfTrials, src/SquareRootCut.cpp:63-83);
 * trials_req > 0: a trial budget with no success guarantee: T = trials_req,
 *   t = min(cac_intermediate_size(n, m), 4096) (the dense engine's fast sizes). */
int cac_serve_mincut_plan(uint32_t n, uint64_t m, double p, uint32_t trials_req, uint32_t* t, uint32_t* trials);
int cac_server_create(int device, uint32_t workers, uint64_t hbm_budget, uint32_t max_batch, cac_server** out);
int cac_server_submit(cac_server* srv, const cac_query* q, uint64_t* id_out);
/* Moves up to cap finished results to out; with wait != 0 blocks until at least
 * one is ready or nothing is pending or running. */
int cac_server_poll(cac_server* srv, cac_query_result* out, uint32_t cap, uint32_t* count_out, int wait);
double cac_server_clock(cac_server* srv);
/* Peak admitted footprint (bytes of device memory the admitted groups were
 * estimated to hold at once) and the number of launch groups so far. */
int cac_server_stats(cac_server* srv, uint64_t* peak_inflight_bytes, uint64_t* groups);
/* Submits q[i] at arrival_s[i] seconds after the call (open-loop replay of an
 * arrival schedule, non-decreasing times) and returns every result, out[i]
 * for q[i].  No other thread may submit meanwhile. */
int cac_server_replay(cac_server* srv, const cac_query* q, const double* arrival_s, uint32_t count,
                      cac_query_result* out);
int cac_server_destroy(cac_server* srv);
/* Device-resident graphs: the edges (cac_edge records, ids < n) are copied to
 * device memory the server holds until unregistered, counted against
 * hbm_budget, and checked once (CAC_EINVAL for ids >= n, CAC_ERANGE for weights
 * >= 2^40 or a sum >= 2^61).  A query with cac_query.graph = *graph_out runs on
 * them with no upload and no per-query check; its results equal those of the
 * same query bringing the edges.  Unregistering a graph that queries still use
 * frees it when the last of them finishes. */
int cac_server_register_graph(cac_server* srv, const cac_edge* edges, uint64_t m, uint32_t n, uint64_t* graph_out);
int cac_server_unregister_graph(cac_server* srv, uint64_t graph);

/* ---- approximate minimum cut ----------------------------------------------
 * Replaces ApproximateCut::run (src/ApproximateCut.cpp:10-79) and numberOfTrials
 * (:82-87).  For i = 1 .. ceil(ln W) (W = total weight): T = cac_approx_trials(p)
 * disjoint copies each keep edge e iff binomial(w_e, 2^-i) > 0, decided exactly as
 * x >= Q_i(w_e) with x = lo64 of Philox-4x32-10(ctr = {i, copy, e_lo, e_hi},
 * key = {seed, 'APRX'}) and Q_i(w) = (1 - 2^-i)^w in 0.64 fixed point (square
 * and multiply, each product floor(a*b / 2^64)); connected components of the
 * union; stop at the first i with more than T components.
 * *approx = 2^i (1 if no level disconnects), *trials = T * i.  n * T < 2^32. */
uint32_t cac_approx_trials(double p);
int cac_approx_cut(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                   uint64_t* approx, uint32_t* trials);

/* ---- synthetic inputs ----------------------------------------------------
 * R-MAT edge list on the device (the reference generates R-MAT inputs with
 * PaRMAT, input_generators/rmat_driver.sh:13-16).  Draw d makes 'scale' quadrant
 * choices from Philox-4x32-10(ctr = {d_lo, level/4, d_hi, 0}, key = {seed, 'RMAT'}),
 * word level%4, against the cumulative u32 thresholds ta < tab < tabc.  Loops
 * are dropped, pairs canonicalised to (min, max), the first m distinct pairs in
 * draw order kept, and d_edges (m entries) is written sorted by (u, v). */
int cac_gen_rmat_dev(cac_ctx* ctx, uint32_t scale, uint64_t m, uint32_t seed, uint32_t ta,
                     uint32_t tab, uint32_t tabc, cac_uedge* d_edges, uint64_t* draws_used);

/* Dense complete graph on the device (the dense Karger-Stein workload; the
 * reference's complete_graph_generator.cpp writes such graphs as text):
 * w(u, v) = 1 + floor(wmax * x / 2^32), x = word 0 of Philox-4x32-10(ctr =
 * {p_lo, p_hi, 0, 0}, key = {seed, 'GENR'}), p = u * n + v for u < v; d_adj is n x n
 * int64, symmetric, zero diagonal. */
int cac_gen_complete_dev(cac_ctx* ctx, uint32_t n, uint32_t seed, uint32_t wmax, int64_t* d_adj);
/* Erdos-Renyi G(n, threshold / 2^32) edge list on the host (the reference's
 * er_generator.cpp): pair (u < v) kept iff word 0 of the same Philox draw <
 * threshold, weight 1 + floor(wmax * word1 / 2^32).  Writes up to cap edges
 * sorted by (u, v); *m_out = edges generated (CAC_ERANGE if more than cap). */
int cac_gen_er(uint32_t n, uint32_t threshold, uint32_t seed, uint32_t wmax, cac_edge* out, uint64_t cap,
               uint64_t* m_out);

/* Planted-cut graphs for the batched min cut (no counterpart in the reference).
 * Graph g (g = first .. first + count - 1) has n_g = 2 floor(e^x / 2) vertices,
 * x uniform in [ln nmin, ln nmax] (word 0 of Philox(ctr = {g, ~0, 0, 0},
 * key = {seed, 'PLNT'})), clamped to even values in [nmin, nmax].  Its halves
 * [0, n/2) and [n/2, n) are G(n/2, avg_degree / (n/2 - 1)) drawn by geometric
 * skipping, with weights uniform in [1, wmax].  c uniform in [1, cmax] distinct
 * unit-weight edges (x < n/2 <= y) join them.  A graph is redrawn until both
 * halves are connected and every weighted degree exceeds c, so c is the
 * weight of the cut between the halves and below every single-vertex cut.
 * cac_gen_planted_sizes gives n_g and an upper bound on the total edge count.
 * cac_gen_planted writes descs[i] (edge_offset, m, n; trials = 0),
 * planted[i] = c_i and the edges of all graphs back to back (*m_total edges;
 * CAC_ERANGE if more than edge_cap).  Host-side, multithreaded. */
/* G(n, avg_degree / (n - 1)) by geometric skipping (no counterpart in the
 * reference): row u draws skips from word 0 of Philox(ctr = {stream, u, j, 0},
 * key = {seed, 'GNPG'}) and weights 1 + floor(wmax * word1 / 2^32); edges u < v
 * sorted by (u, v), loops and duplicates impossible, connectivity not enforced.
 * Writes up to cap edges (*m_out = edges generated; CAC_ERANGE if more).
 * Host-side, single-threaded, reentrant. */
int cac_gen_gnp(uint32_t n, double avg_degree, uint32_t seed, uint32_t stream, uint32_t wmax, cac_edge* out,
                uint64_t cap, uint64_t* m_out);
/* Watts-Strogatz small-world graph, the shape of the reference's sw_generator
 * (src/executables/sw_generator.cpp: Boost small_world_iterator with k = 2 d
 * neighbours, rewiring probability 0.3, every weight 100), from this library's
 * own draws: vertex u links to u + j mod n for j = 1..d, and that edge is
 * rewired with probability beta to (u, x), x uniform over the other n - 1
 * vertices; edge (u, j) uses Philox-4x32-10(ctr = {u, j, 0, 0},
 * key = {seed, 'SWRL'}): rewired iff word 0 < round(beta 2^32), x = u + 1 +
 * floor((n - 1) word1 / 2^32) mod n.  Writes m = n d edges (min, max, weight)
 * in (u, j) order; parallel edges possible.  out = NULL only reports *m_out.
 * Needs n > 2 d.  Host-side. */
int cac_gen_ws(uint32_t n, uint32_t d, double beta, uint32_t seed, uint64_t weight, cac_edge* out, uint64_t cap,
               uint64_t* m_out);
int cac_gen_planted_sizes(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax,
                          double avg_degree, uint32_t cmax, uint32_t* n_out, uint64_t* edge_bound);
int cac_gen_planted(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax, double avg_degree,
                    uint32_t wmax, uint32_t cmax, cac_graph_desc* descs, cac_edge* edges, uint64_t edge_cap,
                    uint64_t* m_total, uint32_t* planted);
/* The hard planted family (no counterpart in the reference): the sizes n_g of
 * cac_gen_planted, but each half is the union of `cycles` Hamiltonian cycles
 * over its n/2 vertices, cycle i in a Philox-shuffled vertex order with every
 * edge of weight w_i uniform in [1, wmax].  Every vertex of a half has
 * weighted degree delta = 2 sum w_i and every cut inside a half crosses each
 * cycle twice, so a half's minimum cut is delta.  c uniform in [1, cmax]
 * distinct cross edges carry C = floor(cut_frac * min(delta_A, delta_B)) in
 * all (split as evenly as integers allow): the minimum cut is C, within a
 * factor cut_frac of the minimum weighted degree -- many single-vertex cuts
 * only slightly heavier than the planted one.  planted[i] = C; m_g =
 * cycles * n_g + c.  Host-side, multithreaded. */
int cac_gen_planted_hard(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax,
                         uint32_t cycles, uint32_t wmax, uint32_t cmax, double cut_frac, cac_graph_desc* descs,
                         cac_edge* edges, uint64_t edge_cap, uint64_t* m_total, uint32_t* planted);

#ifdef __cplusplus
}
#endif

#endif /* CAC_H */
