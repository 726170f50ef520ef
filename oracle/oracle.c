/*
 * oracle.c -- CPU restatement of the contraction core (TEST INFRASTRUCTURE).
 * See oracle.h: only tests/, smoke() and bench.py's cpu_baseline use this.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as    */
/* 1, 2, 3", SC'11).  Round: multiply the even words by the round multipliers, */
/* cross the high halves with the odd words and the key; Weyl key schedule.    */
/* ------------------------------------------------------------------------- */
#define PHILOX_M4x32_0 0xD2511F53u
#define PHILOX_M4x32_1 0xCD9E8D57u
#define PHILOX_W32_0 0x9E3779B9u
#define PHILOX_W32_1 0xBB67AE85u
#define PHILOX_M4x64_0 0xD2E7470EE14C6C93ull
#define PHILOX_M4x64_1 0xCA5A826395121157ull
#define PHILOX_W64_0 0x9E3779B97F4A7C15ull
#define PHILOX_W64_1 0xBB67AE8584CAA73Bull

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W32_0;
            k1 += PHILOX_W32_1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M4x32_0 * x0;
        uint64_t p1 = (uint64_t)PHILOX_M4x32_1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
    uint64_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W64_0;
            k1 += PHILOX_W64_1;
        }
        __uint128_t p0 = (__uint128_t)PHILOX_M4x64_0 * x0;
        __uint128_t p1 = (__uint128_t)PHILOX_M4x64_1 * x2;
        uint64_t y0 = (uint64_t)(p1 >> 64) ^ x1 ^ k0;
        uint64_t y1 = (uint64_t)p1;
        uint64_t y2 = (uint64_t)(p0 >> 64) ^ x3 ^ k1;
        uint64_t y3 = (uint64_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

uint64_t orc_uniform1(uint64_t x, uint64_t W) {
    return 1u + (uint64_t)(((__uint128_t)x * W) >> 64);
}

/* ------------------------------------------------------------------------- */
/* Connected components: union by size + full path compression, then          */
/* canonicalise each component to its minimum vertex id.                       */
/* ------------------------------------------------------------------------- */
static uint32_t uf_find(uint32_t* p, uint32_t x) {
    uint32_t r = x;
    while (p[r] != r) r = p[r];
    while (p[x] != r) {
        uint32_t nx = p[x];
        p[x] = r;
        x = nx;
    }
    return r;
}

uint32_t orc_cc(uint32_t n, uint64_t m, const uint32_t* edges, int mode, uint32_t* labels) {
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t* sz = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) { p[i] = i; sz[i] = 1; }
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a = uf_find(p, edges[2 * e]), b = uf_find(p, edges[2 * e + 1]);
        if (a == b) continue;
        if (sz[a] < sz[b]) { uint32_t t = a; a = b; b = t; }
        p[b] = a;
        sz[a] += sz[b];
    }
    /* minimum id per root: vertices visited in increasing order, so the first
     * vertex seen for a root is its minimum */
    uint32_t* minid = sz; /* reuse */
    for (uint32_t i = 0; i < n; ++i) minid[i] = UINT32_MAX;
    uint32_t ncomp = 0;
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t r = uf_find(p, v);
        if (minid[r] == UINT32_MAX) {
            minid[r] = (mode == 1) ? ncomp : v;
            ++ncomp;
        }
        labels[v] = minid[r];
    }
    free(p);
    free(sz);
    return ncomp;
}

/* ------------------------------------------------------------------------- */
/* Connected components on several threads (the multithreaded CPU baseline):   */
/* the edge list is cut into one contiguous chunk per thread; each thread      */
/* unions its edges with CAS hooking of the larger root under the smaller      */
/* (path halving on every find), so every root is the minimum id of its tree.  */
/* After the join, labels = roots on all threads, then the dense rank (mode 1) */
/* in one sequential pass.  Same output as orc_cc.  The reference spreads the  */
/* same edge-list union-find over MPI ranks (src/executables/parallel_cc.cpp,  */
/* src/UnweightedIteratedSparseSampling.cpp:40-72); here threads share memory.  */
/* ------------------------------------------------------------------------- */
#include <pthread.h>
#include <unistd.h>

typedef struct {
    uint32_t* p;
    const uint32_t* edges;
    uint64_t lo, hi; /* edges [lo, hi) or vertices [lo, hi) */
    uint32_t* labels;
} mt_job;

static uint32_t mt_find(uint32_t* p, uint32_t x) {
    for (;;) {
        uint32_t px = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
        if (px == x) return x;
        uint32_t gx = __atomic_load_n(&p[px], __ATOMIC_RELAXED);
        if (gx != px) __atomic_store_n(&p[x], gx, __ATOMIC_RELAXED); /* an ancestor: valid whatever races */
        x = gx;
    }
}

static void* mt_init(void* arg) {
    mt_job* j = (mt_job*)arg;
    for (uint64_t v = j->lo; v < j->hi; ++v) j->p[v] = (uint32_t)v;
    return NULL;
}

static void* mt_union(void* arg) {
    mt_job* j = (mt_job*)arg;
    for (uint64_t e = j->lo; e < j->hi; ++e) {
        uint32_t a = j->edges[2 * e], b = j->edges[2 * e + 1];
        for (;;) {
            a = mt_find(j->p, a);
            b = mt_find(j->p, b);
            if (a == b) break;
            uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            uint32_t expect = hi;
            if (__atomic_compare_exchange_n(&j->p[hi], &expect, lo, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) break;
            a = expect; /* hi was hooked meanwhile: climb from its new parent */
            b = lo;
        }
    }
    return NULL;
}

static void* mt_label(void* arg) {
    mt_job* j = (mt_job*)arg;
    for (uint64_t v = j->lo; v < j->hi; ++v) j->labels[v] = mt_find(j->p, (uint32_t)v);
    return NULL;
}

static void mt_run(void* (*fn)(void*), mt_job* jobs, int T, uint64_t total, uint32_t* p, const uint32_t* edges,
                   uint32_t* labels) {
    pthread_t th[256];
    for (int t = 0; t < T; ++t) {
        jobs[t].p = p;
        jobs[t].edges = edges;
        jobs[t].labels = labels;
        jobs[t].lo = total * (uint64_t)t / (uint64_t)T;
        jobs[t].hi = total * (uint64_t)(t + 1) / (uint64_t)T;
        pthread_create(&th[t], NULL, fn, &jobs[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
}

uint32_t orc_cc_mt(uint32_t n, uint64_t m, const uint32_t* edges, int mode, uint32_t* labels, int threads) {
    int T = threads > 0 ? threads : (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (T < 1) T = 1;
    if (T > 256) T = 256;
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    mt_job jobs[256];
    mt_run(mt_init, jobs, T, n, p, edges, labels);
    mt_run(mt_union, jobs, T, m, p, edges, labels);
    mt_run(mt_label, jobs, T, n, p, edges, labels);
    uint32_t ncomp = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (labels[v] == v) {
            p[v] = ncomp; /* p now maps a root to its rank */
            ++ncomp;
        }
    }
    if (mode == 1)
        for (uint32_t v = 0; v < n; ++v) labels[v] = p[labels[v]];
    free(p);
    return ncomp;
}

/* ------------------------------------------------------------------------- */
/* R-MAT generator                                                            */
/* ------------------------------------------------------------------------- */
#define DOM_RMAT 0x524D4154u /* 'RMAT' */

static void rmat_draw(uint32_t scale, uint64_t d, uint32_t seed, uint32_t ta, uint32_t tab,
                      uint32_t tabc, uint32_t* u_out, uint32_t* v_out) {
    uint32_t u = 0, v = 0;
    uint32_t key[2] = {seed, DOM_RMAT};
    uint32_t word[4] = {0, 0, 0, 0};
    for (uint32_t lvl = 0; lvl < scale; ++lvl) {
        if ((lvl & 3u) == 0) {
            uint32_t ctr[4] = {(uint32_t)d, lvl >> 2, (uint32_t)(d >> 32), 0};
            orc_philox4x32_10(ctr, key, word);
        }
        uint32_t r = word[lvl & 3u];
        uint32_t bu, bv;
        if (r < ta) { bu = 0; bv = 0; }
        else if (r < tab) { bu = 0; bv = 1; }
        else if (r < tabc) { bu = 1; bv = 0; }
        else { bu = 1; bv = 1; }
        u = (u << 1) | bu;
        v = (v << 1) | bv;
    }
    *u_out = u;
    *v_out = v;
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

uint64_t orc_rmat(uint32_t scale, uint64_t m, uint32_t seed, uint32_t ta, uint32_t tab,
                  uint32_t tabc, uint32_t* edges_out) {
    /* open-addressing hash set of canonical keys */
    uint64_t cap = 16;
    while (cap < 4 * m) cap <<= 1;
    uint64_t* table = (uint64_t*)malloc(sizeof(uint64_t) * cap);
    for (uint64_t i = 0; i < cap; ++i) table[i] = UINT64_MAX;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    uint64_t have = 0, d = 0;
    while (have < m) {
        uint32_t u, v;
        rmat_draw(scale, d, seed, ta, tab, tabc, &u, &v);
        ++d;
        if (u == v) continue;
        uint64_t k = u < v ? (((uint64_t)u << 32) | v) : (((uint64_t)v << 32) | u);
        uint64_t h = (k * 0x9E3779B97F4A7C15ull) >> 17;
        for (;;) {
            uint64_t slot = h & (cap - 1);
            if (table[slot] == UINT64_MAX) {
                table[slot] = k;
                keys[have++] = k;
                break;
            }
            if (table[slot] == k) break;
            ++h;
        }
    }
    qsort(keys, m, sizeof(uint64_t), cmp_u64);
    for (uint64_t i = 0; i < m; ++i) {
        edges_out[2 * i] = (uint32_t)(keys[i] >> 32);
        edges_out[2 * i + 1] = (uint32_t)keys[i];
    }
    free(keys);
    free(table);
    return d;
}

/* ------------------------------------------------------------------------- */
/* Exact minimum cut                                                          */
/* ------------------------------------------------------------------------- */
int64_t orc_stoer_wagner(uint32_t n, const int64_t* adj) {
    if (n < 2) return 0;
    int64_t* w = (int64_t*)malloc(sizeof(int64_t) * (size_t)n * n);
    memcpy(w, adj, sizeof(int64_t) * (size_t)n * n);
    uint32_t* alive = (uint32_t*)malloc(sizeof(uint32_t) * n); /* list of live vertices */
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * n);
    unsigned char* added = (unsigned char*)malloc(n);
    uint32_t nl = n;
    for (uint32_t i = 0; i < n; ++i) alive[i] = i;
    int64_t best = INT64_MAX;
    while (nl > 1) {
        /* maximum adjacency ordering from alive[0] */
        for (uint32_t i = 0; i < nl; ++i) { key[alive[i]] = 0; added[alive[i]] = 0; }
        uint32_t prev = alive[0], last = alive[0];
        added[last] = 1;
        for (uint32_t i = 0; i < nl; ++i) key[alive[i]] += w[(size_t)last * n + alive[i]];
        for (uint32_t step = 1; step < nl; ++step) {
            uint32_t sel = UINT32_MAX;
            for (uint32_t i = 0; i < nl; ++i) {
                uint32_t x = alive[i];
                if (!added[x] && (sel == UINT32_MAX || key[x] > key[sel])) sel = x;
            }
            prev = last;
            last = sel;
            added[sel] = 1;
            if (step == nl - 1) break;
            for (uint32_t i = 0; i < nl; ++i) {
                uint32_t x = alive[i];
                if (!added[x]) key[x] += w[(size_t)sel * n + x];
            }
        }
        /* cut of the phase: weight of 'last' to everything else */
        int64_t cut = 0;
        for (uint32_t i = 0; i < nl; ++i) cut += w[(size_t)last * n + alive[i]];
        if (cut < best) best = cut;
        /* merge last into prev */
        for (uint32_t i = 0; i < nl; ++i) {
            uint32_t x = alive[i];
            if (x == prev || x == last) continue;
            w[(size_t)prev * n + x] += w[(size_t)last * n + x];
            w[(size_t)x * n + prev] = w[(size_t)prev * n + x];
        }
        w[(size_t)prev * n + last] = w[(size_t)last * n + prev] = 0;
        uint32_t j = 0;
        for (uint32_t i = 0; i < nl; ++i)
            if (alive[i] != last) alive[j++] = alive[i];
        nl = j;
    }
    free(w);
    free(alive);
    free(key);
    free(added);
    return best;
}

int64_t orc_brute_mincut(uint32_t n, const int64_t* adj) {
    if (n < 2 || n > 24) return -1;
    /* S ranges over nonempty subsets of {0..n-2}; vertex n-1 is always outside.
     * Gray-code walk: flipping vertex x changes the cut by
     * (weight from x to the other side) - (weight from x to its own side). */
    uint32_t k = n - 1;
    unsigned char* in = (unsigned char*)calloc(n, 1);
    int64_t cut = 0, best = INT64_MAX;
    for (uint64_t g = 1; g < (1ull << k); ++g) {
        uint32_t x = (uint32_t)__builtin_ctzll(g);
        int64_t same = 0, other = 0;
        for (uint32_t y = 0; y < n; ++y) {
            if (y == x) continue;
            if (in[y] == in[x]) same += adj[(size_t)x * n + y];
            else other += adj[(size_t)x * n + y];
        }
        /* moving x to the other side: edges to its current side become cut,
         * edges to the other side stop being cut */
        cut += same - other;
        in[x] ^= 1;
        if (cut < best) best = cut;
    }
    free(in);
    return best;
}

/* ------------------------------------------------------------------------- */
/* Trial-count and size rules                                                */
/* ------------------------------------------------------------------------- */
#include <math.h>

#define KS_BASE 128u
#define DOM_KS 0x4B415247u     /* 'KARG' */
#define DOM_SAMPLE 0x53414D50u /* 'SAMP' */

uint32_t orc_ks_target(uint32_t s) {
    int t = (int)ceil((double)s / sqrt((double)2)) + 1;
    return t > (int)KS_BASE ? (uint32_t)t : KS_BASE;
}

double orc_ks_min_success(uint32_t s) {
    if (s <= KS_BASE) return 1.0;
    double r = orc_ks_min_success(orc_ks_target(s));
    return 1.0 - pow((1 - 0.5 * r), 2);
}

uint32_t orc_ks_trials(uint32_t n, double p) {
    if (n <= KS_BASE) return 1;
    if (p == 0) return 1;
    double s = orc_ks_min_success(n), cur = 1.0, target = 1.0 - p;
    uint32_t k = 0;
    while (cur > target) {
        cur = cur * (1.0 - s);
        ++k;
    }
    return k;
}

uint32_t orc_intermediate_size(uint32_t n, uint64_t m) {
    uint32_t c = (uint32_t)ceil(2.0 * sqrt((double)(uint32_t)m) + 1);
    return c > n ? n : c;
}

uint32_t orc_sqrt_trials(uint32_t n, uint64_t m, double p) {
    double pk = orc_ks_min_success(orc_intermediate_size(n, m));
    double cp = 1.0 / (1 - p);
    return (uint32_t)ceil((n * (n * log(cp))) / (2.0 * 2.0 * ((uint32_t)m * pk)));
}

uint32_t orc_approx_trials(double p) {
    double alpha = log(1 - exp(-(double)0.75));
    return (uint32_t)ceil(log((double)1.0 - p) / alpha);
}

/* ------------------------------------------------------------------------- */
/* Dense Karger-Stein                                                        */
/* ------------------------------------------------------------------------- */
int64_t orc_min_degree(uint32_t n, const int64_t* adj) {
    int64_t best = INT64_MAX;
    for (uint32_t i = 0; i < n; ++i) {
        int64_t d = 0;
        for (uint32_t j = 0; j < n; ++j) d += adj[(size_t)i * n + j];
        if (d < best) best = d;
    }
    return n ? best : 0;
}

typedef struct {
    uint32_t s;
    int64_t* A;     /* s x s working copy */
    uint32_t* rep;  /* representative of every original row */
    int64_t* cap;   /* weighted degree of every live row, 0 for dead rows */
    int64_t W;      /* sum of cap */
    uint32_t nv;    /* live rows */
} ks_state;

/* lazy_adjacency_matrix::contract_edge + indexed capacity update: row a dies into row b */
static int64_t ks_merge(ks_state* g, uint32_t a, uint32_t b) {
    uint32_t s = g->s;
    for (uint32_t x = 0; x < s; ++x)
        if (g->rep[x] == a) g->rep[x] = b;
    int64_t* ra = g->A + (size_t)a * s;
    int64_t* rb = g->A + (size_t)b * s;
    int64_t sum = 0;
    for (uint32_t x = 0; x < s; ++x) {
        int64_t v = rb[x] + ra[x];
        if (g->rep[x] == b) v = 0;
        rb[x] = v;
        sum += v;
    }
    int64_t w = (g->cap[a] + g->cap[b] - sum) / 2;
    g->W -= 2 * w;
    g->cap[b] = sum;
    g->cap[a] = 0;
    g->nv -= 1;
    return w;
}

static uint32_t first_prefix_ge(const int64_t* v, uint32_t len, uint64_t r) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < len; ++i) {
        acc += (uint64_t)v[i];
        if (acc >= r) return i;
    }
    return len; /* unreachable for 1 <= r <= sum */
}

/* contract a copy of A (s x s) to tgt super-vertices; returns the compacted
 * k x k matrix (malloc'd) or NULL when the graph ran out of edges (disconnected). */
static int64_t* ks_contract(uint32_t s, const int64_t* A, uint32_t tgt, uint32_t seed, uint32_t trial,
                            uint32_t nid, int64_t bound, uint32_t* k_out) {
    ks_state g;
    g.s = s;
    g.A = (int64_t*)malloc(sizeof(int64_t) * (size_t)s * s);
    memcpy(g.A, A, sizeof(int64_t) * (size_t)s * s);
    g.rep = (uint32_t*)malloc(sizeof(uint32_t) * s);
    g.cap = (int64_t*)malloc(sizeof(int64_t) * s);
    g.W = 0;
    g.nv = s;
    for (uint32_t x = 0; x < s; ++x) {
        g.rep[x] = x;
        int64_t d = 0;
        for (uint32_t y = 0; y < s; ++y) d += A[(size_t)x * s + y];
        g.cap[x] = d;
        g.W += d;
    }
    uint32_t key[2] = {seed, DOM_KS};
    int pr1 = 0, disconnected = 0;
    for (uint32_t p = 0; p + tgt < s; ++p) {
        if (g.W == 0) {
            disconnected = 1;
            break;
        }
        uint32_t ctr[4] = {trial, nid, p, 0}, o[4];
        orc_philox4x32_10(ctr, key, o);
        uint64_t xr = ((uint64_t)o[1] << 32) | o[0];
        uint64_t xc = ((uint64_t)o[3] << 32) | o[2];
        uint32_t i = first_prefix_ge(g.cap, s, orc_uniform1(xr, (uint64_t)g.W));
        uint32_t k = first_prefix_ge(g.A + (size_t)i * s, s, orc_uniform1(xc, (uint64_t)g.cap[i]));
        uint32_t j = g.rep[k];
        uint32_t a = i > j ? i : j, b = i > j ? j : i;
        int64_t w = ks_merge(&g, a, b);
        if (w >= bound) ++pr1;
    }
    int64_t* out = NULL;
    if (!disconnected) {
        /* indexed_adjacency_matrix::pr_pass: contract any edge of weight >= bound */
        if (pr1 > 0 && bound > 0) {
            for (uint32_t i = 0; i < s && g.nv > 2; ++i) {
                if (g.rep[i] != i || g.cap[i] < bound) continue;
                const int64_t* row = g.A + (size_t)i * s;
                for (uint32_t y = 0; y < s; ++y) {
                    if (row[y] >= bound) {
                        ks_merge(&g, i, g.rep[y]);
                        break;
                    }
                }
            }
        }
        /* compaction: rank of live rows; sum columns by class */
        uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * s);
        uint32_t k = 0;
        for (uint32_t x = 0; x < s; ++x)
            if (g.rep[x] == x) rank[x] = k++;
        out = (int64_t*)calloc((size_t)k * k, sizeof(int64_t));
        for (uint32_t a = 0; a < s; ++a) {
            if (g.rep[a] != a) continue;
            int64_t* dst = out + (size_t)rank[a] * k;
            const int64_t* row = g.A + (size_t)a * s;
            for (uint32_t y = 0; y < s; ++y) dst[rank[g.rep[y]]] += row[y];
        }
        *k_out = k;
        free(rank);
    }
    free(g.A);
    free(g.rep);
    free(g.cap);
    return out;
}

static int64_t ks_node(uint32_t s, const int64_t* A, uint32_t seed, uint32_t trial, uint32_t nid, int64_t bound) {
    if (s <= KS_BASE) return orc_stoer_wagner(s, A);
    uint32_t tgt = orc_ks_target(s);
    int64_t best = INT64_MAX;
    for (uint32_t c = 0; c < 2; ++c) {
        uint32_t cid = 2 * nid + c, k = 0;
        int64_t* C = ks_contract(s, A, tgt, seed, trial, cid, bound, &k);
        int64_t v = C ? ks_node(k, C, seed, trial, cid, bound) : 0;
        free(C);
        if (v < best) best = v;
    }
    return best;
}

int64_t orc_ks_trial(uint32_t n, const int64_t* adj, uint32_t seed, uint32_t trial) {
    if (n < 2) return 0;
    int64_t bound = orc_min_degree(n, adj);
    int64_t v = ks_node(n, adj, seed, trial, 1, bound);
    return v < bound ? v : bound;
}

/* ------------------------------------------------------------------------- */
/* Weighted iterated sampling                                                */
/* ------------------------------------------------------------------------- */
static uint32_t root_min(uint32_t* p, uint32_t x) {
    uint32_t r = x;
    while (p[r] != r) r = p[r];
    while (p[x] != r) {
        uint32_t nx = p[x];
        p[x] = r;
        x = nx;
    }
    return r;
}

/* Draws per round of the high-concurrency weighted sampling:
 * WeightedIteratedSparseSampling::edgesToSamplePerProcessor
 * (src/WeightedIteratedSparseSampling.hpp:66-84) draws (unsigned)
 * pow((float) n0, 1 + epsilon / 2) edges per round, epsilon = 0.1f
 * (src/IteratedSparseSampling.hpp:38), n0 the initial vertex count, all in float. */
uint32_t orc_hc_draws(uint32_t n0) {
    const float epsilon = 0.1f;
    return (uint32_t)powf((float)n0, 1 + epsilon / 2);
}

uint32_t orc_sample_contract(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed,
                             uint32_t trial, uint32_t* labels) {
    return orc_sample_contract_rule(n, m, edges, t, seed, trial, 0, labels);
}

/* The round loop of both sampling rules: a round draws draws_per_round edges,
 * or |V| when draws_per_round == 0 (the low-concurrency rule,
 * src/SequentialSquareRootCut.cpp:37-48); the high-concurrency rule draws
 * orc_hc_draws(n) per round over the current graph
 * (src/WeightedIteratedSparseSampling.cpp:236-279) and contracts a draw prefix
 * to the target (src/IteratedSparseSampling.hpp:182-224); its gathered draws
 * are shuffled (src/IteratedSparseSampling.hpp:278), which leaves i.i.d. draws
 * i.i.d., so they are used in draw order. */
uint32_t orc_sample_contract_rule(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed,
                                  uint32_t trial, uint32_t draws_per_round, uint32_t* labels) {
    uint64_t* P = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    uint32_t* par = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t key[2] = {seed, DOM_SAMPLE};
    for (uint32_t v = 0; v < n; ++v) labels[v] = v;
    uint32_t nv = n;
    for (uint32_t round = 0; nv > t; ++round) {
        uint64_t W = 0;
        for (uint64_t e = 0; e < m; ++e) {
            uint32_t u = (uint32_t)edges[3 * e], v = (uint32_t)edges[3 * e + 1];
            if (labels[u] != labels[v]) W += edges[3 * e + 2];
            P[e] = W;
        }
        if (W == 0) break;
        for (uint32_t v = 0; v < n; ++v) par[v] = v;
        uint32_t draws = draws_per_round ? draws_per_round : nv;
        for (uint32_t d = 0; d < draws; ++d) {
            uint32_t ctr[4] = {trial, round, d, 0}, o[4];
            orc_philox4x32_10(ctr, key, o);
            uint64_t r = orc_uniform1(((uint64_t)o[1] << 32) | o[0], W);
            uint64_t lo = 0, hi = m - 1; /* first e with P[e] >= r */
            while (lo < hi) {
                uint64_t mid = (lo + hi) / 2;
                if (P[mid] >= r) hi = mid;
                else lo = mid + 1;
            }
            uint32_t a = root_min(par, labels[edges[3 * lo]]);
            uint32_t b = root_min(par, labels[edges[3 * lo + 1]]);
            if (a == b) continue;
            if (a < b) par[b] = a;
            else par[a] = b;
            if (--nv == t) break;
        }
        for (uint32_t v = 0; v < n; ++v) labels[v] = root_min(par, labels[v]);
    }
    free(P);
    free(par);
    return nv;
}

uint32_t orc_densify(uint32_t n, uint64_t m, const uint64_t* edges, const uint32_t* labels, int64_t* out) {
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t k = 0;
    for (uint32_t x = 0; x < n; ++x)
        if (labels[x] == x) rank[x] = k++;
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a = rank[labels[edges[3 * e]]], b = rank[labels[edges[3 * e + 1]]];
        if (a == b) continue;
        out[(size_t)a * k + b] += (int64_t)edges[3 * e + 2];
        out[(size_t)b * k + a] += (int64_t)edges[3 * e + 2];
    }
    free(rank);
    return k;
}

#define SQRT_DENSE_MAX 8192u /* the dense engine's largest matrix */
int64_t orc_ks_sparse_trial(uint32_t n, uint64_t m, const uint64_t* E, uint32_t seed, uint32_t trial);

static int cmp_u64x3(const void* a, const void* b) {
    const uint64_t* x = (const uint64_t*)a;
    const uint64_t* y = (const uint64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

/* The contracted graph as an edge list, the form SequentialKargerSteinCut gets
 * (AdjacencyListGraph::compact, src/AdjacencyListGraph.cpp:34-112): endpoints
 * ranked by canonical label, loops dropped, parallel edges summed, sorted by
 * (u, v).  Returns the list (caller frees), *m_out entries. */
static uint64_t* merged_contracted(uint32_t n, uint64_t m, const uint64_t* edges, const uint32_t* lab,
                                   uint64_t* m_out) {
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t r = 0;
    for (uint32_t x = 0; x < n; ++x)
        if (lab[x] == x) rank[x] = r++;
    uint64_t* T = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (m ? m : 1));
    uint64_t k = 0;
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a = rank[lab[edges[3 * e]]], b = rank[lab[edges[3 * e + 1]]];
        if (a == b) continue;
        T[3 * k] = a < b ? a : b;
        T[3 * k + 1] = a < b ? b : a;
        T[3 * k + 2] = edges[3 * e + 2];
        ++k;
    }
    qsort(T, k, 3 * sizeof(uint64_t), cmp_u64x3);
    uint64_t out = 0;
    for (uint64_t i = 0; i < k; ++i) {
        if (out > 0 && T[3 * (out - 1)] == T[3 * i] && T[3 * (out - 1) + 1] == T[3 * i + 1]) {
            T[3 * (out - 1) + 2] += T[3 * i + 2];
        } else {
            T[3 * out] = T[3 * i];
            T[3 * out + 1] = T[3 * i + 1];
            T[3 * out + 2] = T[3 * i + 2];
            ++out;
        }
    }
    free(rank);
    *m_out = out;
    return T;
}

int64_t orc_sqrt_trial(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed, uint32_t trial) {
    return orc_sqrt_trial_rule(n, m, edges, t, seed, trial, 0);
}

int64_t orc_sqrt_trial_rule(uint32_t n, uint64_t m, const uint64_t* edges, uint32_t t, uint32_t seed, uint32_t trial,
                            uint32_t draws_per_round) {
    uint32_t* lab = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    orc_sample_contract_rule(n, m, edges, t, seed, trial, draws_per_round, lab);
    uint32_t k = 0;
    for (uint32_t x = 0; x < n; ++x) k += lab[x] == x;
    if (t > SQRT_DENSE_MAX) {
        /* beyond the dense engine: the merged contracted edge list through the
         * sparse recursion (trial `trial`, node 1); disconnected: cut 0 */
        int64_t cut = 0;
        if (k <= t) {
            uint64_t mm = 0;
            uint64_t* E = merged_contracted(n, m, edges, lab, &mm);
            cut = orc_ks_sparse_trial(k, mm, E, seed, trial);
            free(E);
        }
        free(lab);
        return cut;
    }
    int64_t* M = (int64_t*)calloc((size_t)k * k, sizeof(int64_t));
    orc_densify(n, m, edges, lab, M);
    int64_t cut = orc_ks_trial(k, M, seed, trial);
    free(M);
    free(lab);
    return cut;
}

/* First pick of node nid on matrix A (the selection rule alone), for the
 * distribution test: pair {a, b} must be drawn with probability A[a][b] / W. */
void orc_ks_first_pick(uint32_t n, const int64_t* adj, uint32_t seed, uint32_t trial, uint32_t nid, uint32_t* a,
                       uint32_t* b) {
    int64_t* cap = (int64_t*)malloc(sizeof(int64_t) * n);
    uint64_t W = 0;
    for (uint32_t x = 0; x < n; ++x) {
        int64_t d = 0;
        for (uint32_t y = 0; y < n; ++y) d += adj[(size_t)x * n + y];
        cap[x] = d;
        W += (uint64_t)d;
    }
    uint32_t key[2] = {seed, DOM_KS}, ctr[4] = {trial, nid, 0, 0}, o[4];
    orc_philox4x32_10(ctr, key, o);
    uint32_t i = first_prefix_ge(cap, n, orc_uniform1(((uint64_t)o[1] << 32) | o[0], W));
    uint32_t k = first_prefix_ge(adj + (size_t)i * n, n, orc_uniform1(((uint64_t)o[3] << 32) | o[2], (uint64_t)cap[i]));
    *a = i > k ? i : k;
    *b = i > k ? k : i;
    free(cap);
}

/* ------------------------------------------------------------------------- */
/* Approximate minimum cut                                                   */
/* ------------------------------------------------------------------------- */
#define DOM_APPROX 0x41505258u /* 'APRX' */

static uint64_t mulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((__uint128_t)a * b) >> 64); }

/* (1 - 2^-i)^w as a 0.64 fixed-point fraction; UINT64_MAX stands for 1 (w == 0). */
uint64_t orc_keep_threshold(uint32_t i, uint64_t w) {
    if (w == 0) return UINT64_MAX;
    uint64_t base = (i >= 64) ? UINT64_MAX : (uint64_t)0 - ((uint64_t)1 << (64 - i));
    int have = 0;
    uint64_t acc = 0;
    while (w) {
        if (w & 1) {
            acc = have ? mulhi64(acc, base) : base;
            have = 1;
        }
        w >>= 1;
        if (w) base = mulhi64(base, base);
    }
    return acc;
}

void orc_approx_cut(uint32_t n, uint64_t m, const uint64_t* edges, double p, uint32_t seed, uint64_t* approx,
                    uint32_t* trials) {
    uint64_t W = 0;
    for (uint64_t e = 0; e < m; ++e) W += edges[3 * e + 2];
    uint32_t iters = (uint32_t)ceil(log((double)W));
    uint32_t T = orc_approx_trials(p);
    uint32_t first = 0;
    uint64_t N = (uint64_t)n * T;
    uint32_t* lab = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t* sub = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (m ? m : 1) * T);
    uint32_t key[2] = {seed, DOM_APPROX};
    for (uint32_t i = 1; i <= iters; ++i) {
        uint64_t k = 0;
        for (uint32_t j = 0; j < T; ++j) {
            for (uint64_t e = 0; e < m; ++e) {
                uint32_t ctr[4] = {i, j, (uint32_t)e, (uint32_t)(e >> 32)}, o[4];
                orc_philox4x32_10(ctr, key, o);
                uint64_t x = ((uint64_t)o[1] << 32) | o[0];
                uint64_t Q = orc_keep_threshold(i, edges[3 * e + 2]);
                if (edges[3 * e + 2] == 0 || x < Q) continue;
                sub[2 * k] = (uint32_t)edges[3 * e] + j * n;
                sub[2 * k + 1] = (uint32_t)edges[3 * e + 1] + j * n;
                ++k;
            }
        }
        uint32_t comps = orc_cc((uint32_t)N, k, sub, 0, lab);
        if (comps > T) {
            first = i;
            break;
        }
    }
    *approx = (uint64_t)1 << first;
    *trials = T * first;
    free(lab);
    free(sub);
}

/* ------------------------------------------------------------------------- */
/* Sparse Karger-Stein (co_mincut.h:74-155, recursion :176-223, 280-293)     */
/* ------------------------------------------------------------------------- */
#define DOM_SPARSE 0x53505253u /* 'SPRS' */
#define SPARSE_DENSE_MAX 2048u

/* Exponential-race key of an edge of weight w from a 64-bit word x:
 * floor(2^20 * L / w), L = -log2(U / 2^53) in Q32.32, U = (x >> 11) + 1.
 * log2 by the classical bit-by-bit squaring of the top 32 mantissa bits (32
 * steps, each square truncated to 32 bits; |error| < 2^-29), so CPU and GPU
 * agree exactly.  Ordering edges by (key, index) is, up to the fixed point,
 * ordering by Exp(w) variables: the exponential race of co_mincut.h:141-144. */
uint64_t orc_race_key(uint64_t x, uint64_t w) {
    uint64_t U = (x >> 11) + 1;
    int e = 63 - __builtin_clzll(U);
    uint32_t y = (uint32_t)((U << (63 - e)) >> 32); /* mantissa in [1, 2), Q1.31 */
    uint64_t frac = 0;
    for (int i = 0; i < 32; ++i) {
        uint64_t sq = (uint64_t)y * y; /* Q2.62 */
        frac <<= 1;
        if (sq >> 63) {
            frac |= 1;
            y = (uint32_t)(sq >> 32);
        } else {
            y = (uint32_t)(sq >> 31);
        }
    }
    uint64_t log2u = ((uint64_t)e << 32) | frac;
    uint64_t L = ((uint64_t)53 << 32) - log2u;
    return (L << 20) / (w ? w : 1);
}

typedef struct {
    uint64_t key;
    uint64_t idx;
} race_item;

static int cmp_race(const void* a, const void* b) {
    const race_item* x = (const race_item*)a;
    const race_item* y = (const race_item*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

/* Contract the sparse graph (n, m, E = u,v,w triples, u < v) to tgt vertices:
 * edges in (key, index) order are united until tgt components remain; the
 * child keeps every edge in its original order relabelled by the rank of its
 * components' minimum ids, loops dropped, parallel edges not merged.  Returns
 * the child edge list (caller frees) and *m_out, or NULL when even all edges
 * leave more than tgt components (disconnected: cut 0). */
uint64_t* orc_sparse_contract(uint32_t n, uint64_t m, const uint64_t* E, uint32_t tgt, uint32_t seed, uint32_t trial,
                              uint32_t nid, uint64_t* m_out) {
    race_item* it = (race_item*)malloc(sizeof(race_item) * (m ? m : 1));
    uint32_t key[2] = {seed, DOM_SPARSE};
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t ctr[4] = {trial, nid, (uint32_t)e, (uint32_t)(e >> 32)}, o[4];
        orc_philox4x32_10(ctr, key, o);
        it[e].key = orc_race_key(((uint64_t)o[1] << 32) | o[0], E[3 * e + 2]);
        it[e].idx = e;
    }
    qsort(it, m, sizeof(race_item), cmp_race);
    uint32_t* par = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t v = 0; v < n; ++v) par[v] = v;
    uint32_t comps = n;
    for (uint64_t k = 0; k < m && comps > tgt; ++k) {
        uint64_t e = it[k].idx;
        uint32_t a = root_min(par, (uint32_t)E[3 * e]), b = root_min(par, (uint32_t)E[3 * e + 1]);
        if (a == b) continue;
        if (a < b) par[b] = a;
        else par[a] = b;
        --comps;
    }
    free(it);
    if (comps > tgt) {
        free(par);
        return NULL;
    }
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t r = 0;
    for (uint32_t v = 0; v < n; ++v)
        if (root_min(par, v) == v) rank[v] = r++;
    uint64_t* C = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (m ? m : 1));
    uint64_t mc = 0;
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a = rank[root_min(par, (uint32_t)E[3 * e])], b = rank[root_min(par, (uint32_t)E[3 * e + 1])];
        if (a == b) continue;
        C[3 * mc] = a < b ? a : b;
        C[3 * mc + 1] = a < b ? b : a;
        C[3 * mc + 2] = E[3 * e + 2];
        ++mc;
    }
    free(par);
    free(rank);
    *m_out = mc;
    return C;
}

static int64_t sparse_node(uint32_t n, uint64_t m, const uint64_t* E, uint32_t seed, uint32_t trial, uint32_t nid,
                           int64_t bound) {
    if (n <= SPARSE_DENSE_MAX || (m >= (uint64_t)n * n / 32 && n <= 8192)) {
        int64_t* A = (int64_t*)calloc((size_t)n * n, sizeof(int64_t));
        for (uint64_t e = 0; e < m; ++e) {
            uint32_t a = (uint32_t)E[3 * e], b = (uint32_t)E[3 * e + 1];
            A[(size_t)a * n + b] += (int64_t)E[3 * e + 2];
            A[(size_t)b * n + a] += (int64_t)E[3 * e + 2];
        }
        int64_t v = ks_node(n, A, seed, trial, nid, bound);
        free(A);
        return v;
    }
    uint32_t tgt = orc_ks_target(n);
    int64_t best = INT64_MAX;
    for (uint32_t c = 0; c < 2; ++c) {
        uint32_t cid = 2 * nid + c;
        uint64_t mc = 0;
        uint64_t* C = orc_sparse_contract(n, m, E, tgt, seed, trial, cid, &mc);
        int64_t v = C ? sparse_node(tgt, mc, C, seed, trial, cid, bound) : 0;
        free(C);
        if (v < best) best = v;
    }
    return best;
}

int64_t orc_ks_sparse_trial(uint32_t n, uint64_t m, const uint64_t* E, uint32_t seed, uint32_t trial) {
    if (n < 2) return 0;
    int64_t* deg = (int64_t*)calloc(n, sizeof(int64_t));
    for (uint64_t e = 0; e < m; ++e) {
        deg[E[3 * e]] += (int64_t)E[3 * e + 2];
        deg[E[3 * e + 1]] += (int64_t)E[3 * e + 2];
    }
    int64_t bound = INT64_MAX;
    for (uint32_t v = 0; v < n; ++v) bound = deg[v] < bound ? deg[v] : bound;
    free(deg);
    int64_t v = sparse_node(n, m, E, seed, trial, 1, bound);
    return v < bound ? v : bound;
}
