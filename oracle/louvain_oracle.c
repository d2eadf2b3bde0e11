/*
 * Louvain oracle: a CPU restatement of the reference LouvainCUDA pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is compared
 * against.  Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it; the product (liblouvain_amd / liblouvain_host) never does.
 *
 * It follows the reference's own data flow rather than the engine's design:
 *   - the level edge list is symmetrised into 2E entries (self-loops twice) and
 *     stably sorted by (source, target)              LouvainCUDA/NeighbourCompute.cu:42-118
 *   - degree k_i = sum (w==0?1:w)*(self?2:1), loop_i = weight of the last self
 *     entry, internals = loop, communities = singletons  LouvainCUDA/InitKernel.cu:37-126
 *   - each pass groups the row's (community, w') pairs in ascending community
 *     order (the reference sorts and reduce_by_keys them), scans the groups with
 *     the reference's replace rule, applies the singleton guard, accumulates the
 *     Jacobi deltas and measures Q on the state before the deltas
 *                                                       LouvainCUDA/OneLevelKernel.cu:28-171,248-451,518-954
 *   - renumbering by first occurrence                   LouvainCUDA/RenumberKernel.cu:63-220
 *   - coarsening into lexicographically sorted ordered pairs
 *                                                       LouvainCUDA/InducedGraphKernel.cu:28-113
 *   - the level loop and its termination rule           LouvainCUDA/LouvainCPP.cpp:209-553
 *   - dendrogram flattening                             LouvainCUDA/LouvainCPP.cpp:82-127
 *
 * Intentional differences from the reference (the engine makes the same ones):
 *   D1 weight sums are exact int64 in units of 2^-s (s = smallest exponent that
 *      makes every input weight an integer), not int32 truncations of float sums;
 *   D2 gains are FP64, evaluated as W - ((gamma*deg[c])*f) with f = k_i/(2m),
 *      no FMA; Q is evaluated from the exact aggregates sum(internals) and
 *      sum(deg^2) as sum_int/m - gamma*sum_deg2/(4m*m);
 *   D3 loop/internals are indexed by vertex (the reference indexes them by CSR
 *      row, LouvainCUDA/InitKernel.cu:122-125, which differs once an isolated
 *      vertex precedes a vertex);
 *   D4 empty rows are skipped (no out-of-bounds read, InitKernel.cu:34);
 *   D5 the argmax implements the stated sequential rule (the reference's shuffle
 *      reads exited lanes for rows with fewer than 32 groups);
 *   D7 E == 0 returns the identity partition (node j -> community j).
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_NOMEM (-1)
#define ORC_ERR_INPUT (-2)
#define ORC_ERR_RANGE (-3)
#define ORC_ERR_ARG (-4)

/* weight sums must stay below this many units so that 4*m fits in int64 */
#define ORC_UNIT_LIMIT ((int64_t)1 << 60)

typedef unsigned __int128 u128;

typedef struct {
    int32_t n;
    int64_t e;
    int32_t *es, *et;      /* level edge list, positions */
    int64_t *ew;           /* weights in units */
    int64_t m;
    int64_t *rowptr;       /* n+1; reference CSR incl. self entries */
    int32_t *col;
    int64_t *cw;
    int64_t *k;
    int64_t *loop;
    double q_init;         /* Q of the initial state of this level */
    double lower;
    double q;              /* value oneLevel returns */
    int32_t passes;
    double *pass_q;        /* Q_p, p = 1..passes (index p-1) */
    int32_t **pass_next;   /* next[] of each pass (trace only) */
    int32_t *c_final;      /* assignment after the level */
    int32_t *newc;         /* renumbered community of each position */
    int32_t kcomm;
    int32_t *part_node, *part_comm;
} orc_level;

typedef struct {
    int status;
    int32_t n0;
    int64_t e0;
    int32_t *ids;
    int scale;
    double gamma;
    double min_gain;
    int32_t max_passes;
    int trace;
    int threads;
    int nlevels, cap;
    orc_level *lv;
    int32_t *final_comm;   /* by level-0 position, flattened at the last level */
    double final_q;
} orc_run;

/* ------------------------------------------------------------------------- */

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* ORACLE_PROGRESS=1: one stderr line per stage and pass, so long runs report
 * as they go. */
static int progress_on(void)
{
    const char *p = getenv("ORACLE_PROGRESS");
    return p && p[0] == '1';
}

static void progress(const char *stage, int level, long long a)
{
    if (!progress_on()) return;
    fprintf(stderr, "[oracle] level %d %s %lld\n", level, stage, a);
    fflush(stderr);
}
static void *xmalloc(size_t n) { return malloc(n ? n : 1); }

/* u128 -> double as hi*2^64 + lo, each conversion rounded to nearest. */
static double u128_to_double(u128 x)
{
    double hi = (double)(uint64_t)(x >> 64);
    double lo = (double)(uint64_t)x;
    return hi * 18446744073709551616.0 + lo;
}

static double q_from_aggregates(int64_t sum_int, u128 sum_deg2, int64_t m, double gamma)
{
    if (m <= 0) return 0.0; /* the reference guards links > 0, OneLevelKernel.cu:447 */
    double dm = (double)m;
    double a = (double)sum_int / dm;
    double b = u128_to_double(sum_deg2) / ((4.0 * dm) * dm);
    return a - gamma * b;
}

static double q_of_state(int32_t n, const int64_t *in, const int64_t *deg, int64_t m, double gamma)
{
    int64_t si = 0;
    u128 sd = 0;
    for (int32_t c = 0; c < n; ++c) {
        si += in[c];
        u128 d = (u128)(uint64_t)deg[c];
        sd += d * d;
    }
    return q_from_aggregates(si, sd, m, gamma);
}

/* Smallest s >= 0 with w * 2^s integral; -1 for an invalid weight. */
static int float_unit_exponent(float wf)
{
    double w = (double)wf;
    if (!(w >= 0.0) || isinf(w)) return -1;
    if (w == 0.0) return 0;
    int s = 0;
    while (s <= 160) {
        double v = ldexp(w, s);
        if (v == floor(v)) return s;
        ++s;
    }
    return -1;
}

/* Stable counting sort of idx[0..len) by key[idx[j]] in [0, nkeys). */
static int counting_sort_stable(int64_t len, const int32_t *key, int32_t nkeys, int64_t *idx)
{
    int64_t *cnt = (int64_t *)xcalloc((size_t)nkeys + 1, sizeof(int64_t));
    int64_t *tmp = (int64_t *)xmalloc((size_t)len * sizeof(int64_t));
    if (!cnt || !tmp) { free(cnt); free(tmp); return ORC_ERR_NOMEM; }
    for (int64_t j = 0; j < len; ++j) cnt[key[idx[j]] + 1]++;
    for (int32_t c = 0; c < nkeys; ++c) cnt[c + 1] += cnt[c];
    for (int64_t j = 0; j < len; ++j) tmp[cnt[key[idx[j]]]++] = idx[j];
    memcpy(idx, tmp, (size_t)len * sizeof(int64_t));
    free(cnt);
    free(tmp);
    return ORC_OK;
}

static void level_free(orc_level *L, int trace)
{
    free(L->es); free(L->et); free(L->ew);
    free(L->rowptr); free(L->col); free(L->cw);
    free(L->k); free(L->loop);
    free(L->pass_q);
    if (L->pass_next) {
        for (int32_t p = 0; p < L->passes; ++p) free(L->pass_next[p]);
        free(L->pass_next);
    }
    (void)trace;
    free(L->c_final); free(L->newc);
    free(L->part_node); free(L->part_comm);
    memset(L, 0, sizeof(*L));
}

/* A.1 CSR: m = sum w; 2E entries (s,t,w) then the reversed copies, stably
 * sorted by (source, target).  NeighbourCompute.cu:46,66-68,82,100,113 */
static int build_csr(orc_level *L)
{
    int64_t e = L->e, n2 = 2 * e;
    int32_t n = L->n;
    L->m = 0;
    for (int64_t j = 0; j < e; ++j) L->m += L->ew[j];
    int32_t *src = (int32_t *)xmalloc((size_t)n2 * sizeof(int32_t));
    int32_t *tgt = (int32_t *)xmalloc((size_t)n2 * sizeof(int32_t));
    int64_t *idx = (int64_t *)xmalloc((size_t)n2 * sizeof(int64_t));
    L->rowptr = (int64_t *)xcalloc((size_t)n + 1, sizeof(int64_t));
    L->col = (int32_t *)xmalloc((size_t)n2 * sizeof(int32_t));
    L->cw = (int64_t *)xmalloc((size_t)n2 * sizeof(int64_t));
    if (!src || !tgt || !idx || !L->rowptr || !L->col || !L->cw) {
        free(src); free(tgt); free(idx);
        return ORC_ERR_NOMEM;
    }
    for (int64_t j = 0; j < e; ++j) {
        src[j] = L->es[j]; tgt[j] = L->et[j];
        src[e + j] = L->et[j]; tgt[e + j] = L->es[j];
    }
    for (int64_t j = 0; j < n2; ++j) idx[j] = j;
    /* LSD: stable by target, then stable by source = stable lexicographic sort */
    int rc = counting_sort_stable(n2, tgt, n, idx);
    if (rc == ORC_OK) rc = counting_sort_stable(n2, src, n, idx);
    if (rc != ORC_OK) { free(src); free(tgt); free(idx); return rc; }
    for (int64_t j = 0; j < n2; ++j) {
        int64_t o = idx[j];
        L->col[j] = tgt[o];
        L->cw[j] = L->ew[o < e ? o : o - e];
        L->rowptr[src[o] + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) L->rowptr[i + 1] += L->rowptr[i];
    free(src); free(tgt); free(idx);
    return ORC_OK;
}

/* A.2 init: k_i and loop_i per vertex.  InitKernel.cu:37-77,101-126 */
static int init_vertices(orc_level *L)
{
    int32_t n = L->n;
    L->k = (int64_t *)xcalloc((size_t)n, sizeof(int64_t));
    L->loop = (int64_t *)xcalloc((size_t)n, sizeof(int64_t));
    if (!L->k || !L->loop) return ORC_ERR_NOMEM;
    for (int32_t i = 0; i < n; ++i) {
        int64_t k = 0, loop = 0;
        for (int64_t p = L->rowptr[i]; p < L->rowptr[i + 1]; ++p) {
            int64_t v = L->cw[p] == 0 ? 1 : L->cw[p];
            if (L->col[p] == i) { v *= 2; loop = L->cw[p]; }
            k += v;
        }
        L->k[i] = k;
        L->loop[i] = loop;
    }
    return ORC_OK;
}

typedef struct { int32_t c; int64_t w; } orc_group;

static int cmp_group(const void *a, const void *b)
{
    int32_t x = ((const orc_group *)a)->c, y = ((const orc_group *)b)->c;
    return (x > y) - (x < y);
}

/* A.4 steps 1-3 for one row: groups of (community, sum w') in ascending
 * community order, then the reference's replace rule.
 * OneLevelKernel.cu:43-49 (self weight := 0), :619,644 (sort + reduce_by_key),
 * :265-270,303-316,330,358 (initial best, gain, replace rule). */
static void row_best(const orc_level *L, int32_t i, const int32_t *C, const int64_t *deg,
                     double gamma, int64_t *acc, int64_t *stamp, int64_t token, orc_group *gbuf,
                     int32_t *out_c, int64_t *out_w, int32_t *out_n)
{
    int32_t ng = 0;
    for (int64_t p = L->rowptr[i]; p < L->rowptr[i + 1]; ++p) {
        int32_t j = L->col[p];
        int32_t c = C[j];
        int64_t w = (j == i) ? 0 : L->cw[p];
        if (stamp[c] != token) {   /* first entry of community c in this row */
            stamp[c] = token;
            acc[c] = 0;
            gbuf[ng++].c = c;
        }
        acc[c] += w;
    }
    for (int32_t g = 0; g < ng; ++g) gbuf[g].w = acc[gbuf[g].c];
    qsort(gbuf, (size_t)ng, sizeof(orc_group), cmp_group);
    double f = (double)L->k[i] / (2.0 * (double)L->m);
    int32_t bc = C[i];
    double bg = 0.0;
    int64_t bw = 0;
    for (int32_t g = 0; g < ng; ++g) {
        if (gbuf[g].w == 0) continue;
        int32_t c = gbuf[g].c;
        double gain = (double)gbuf[g].w - ((gamma * (double)deg[c]) * f);
        if (gain > bg || (gain == bg && bg != 0.0 && c < bc)) {
            bc = c;
            bg = gain;
            bw = gbuf[g].w;
        }
    }
    *out_c = bc;
    *out_w = bw;
    *out_n = ng;
}

static void atomic_add64(int64_t *p, int64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
static void atomic_add32(int32_t *p, int32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

/* A.4/A.5: the pass loop and the merge-isolated post-step.
 * OneLevelKernel.cu:518-954 */
static int one_level(orc_run *R, orc_level *L)
{
    int32_t n = L->n;
    double gamma = R->gamma;
    int32_t *C = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int32_t *next = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int64_t *deg = (int64_t *)xmalloc((size_t)n * sizeof(int64_t));
    int64_t *in = (int64_t *)xmalloc((size_t)n * sizeof(int64_t));
    int32_t *size = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int64_t *ddeg = (int64_t *)xcalloc((size_t)n, sizeof(int64_t));
    int64_t *din = (int64_t *)xcalloc((size_t)n, sizeof(int64_t));
    int32_t *dsize = (int32_t *)xcalloc((size_t)n, sizeof(int32_t));
    int32_t *best_c = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int64_t *best_w = (int64_t *)xmalloc((size_t)n * sizeof(int64_t));
    int32_t *ngr = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int cap = 64;
    L->pass_q = (double *)xmalloc((size_t)cap * sizeof(double));
    if (R->trace) L->pass_next = (int32_t **)xcalloc((size_t)cap, sizeof(int32_t *));
    int rc = ORC_OK;
    if (!C || !next || !deg || !in || !size || !ddeg || !din || !dsize || !best_c || !best_w ||
        !ngr || !L->pass_q || (R->trace && !L->pass_next)) { rc = ORC_ERR_NOMEM; goto out; }

    int64_t maxrow = 0;
    for (int32_t i = 0; i < n; ++i) {
        C[i] = i;
        deg[i] = L->k[i];          /* InitKernel.cu:119 */
        in[i] = L->loop[i];        /* InitKernel.cu:125 (D3: by vertex) */
        size[i] = 1;               /* OneLevelKernel.cu:87 */
        int64_t d = L->rowptr[i + 1] - L->rowptr[i];
        if (d > maxrow) maxrow = d;
    }

    int nthreads = 1;
#ifdef _OPENMP
    nthreads = R->threads > 0 ? R->threads : 1;
#endif
    int64_t **accs = (int64_t **)xcalloc((size_t)nthreads, sizeof(int64_t *));
    int64_t **stamps = (int64_t **)xcalloc((size_t)nthreads, sizeof(int64_t *));
    /* one token per (pass, row): stamps never repeat within a level */
    int64_t *tokens = (int64_t *)xcalloc((size_t)nthreads, sizeof(int64_t));
    orc_group **gbufs = (orc_group **)xcalloc((size_t)nthreads, sizeof(orc_group *));
    if (!accs || !stamps || !gbufs || !tokens) {
        free(accs); free(stamps); free(gbufs); free(tokens);
        rc = ORC_ERR_NOMEM;
        goto out;
    }
    for (int t = 0; t < nthreads; ++t) {
        accs[t] = (int64_t *)xmalloc((size_t)n * sizeof(int64_t));
        stamps[t] = (int64_t *)xcalloc((size_t)n, sizeof(int64_t));
        gbufs[t] = (orc_group *)xmalloc((size_t)(maxrow + 1) * sizeof(orc_group));
        if (!accs[t] || !stamps[t] || !gbufs[t]) rc = ORC_ERR_NOMEM;
    }
    if (rc != ORC_OK) goto out_scratch;

    double curQ = -1.0; /* OneLevelKernel.cu:528 */
    int32_t pass = 0;
    for (;;) {
        ++pass;
        if (progress_on()) {
            fprintf(stderr, "[oracle] pass %d of a level with n %d\n", (int)pass, (int)n);
            fflush(stderr);
        }
        /* steps 1-5: every row reads the same state S (Jacobi) */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
#endif
        for (int32_t i = 0; i < n; ++i) {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            if (L->rowptr[i + 1] == L->rowptr[i]) { /* no row: never written */
                ngr[i] = 0;
                next[i] = C[i];
                continue;
            }
            int32_t bc;
            int64_t bw;
            int32_t ng;
            row_best(L, i, C, deg, gamma, accs[t], stamps[t], ++tokens[t], gbufs[t], &bc, &bw, &ng);
            best_c[i] = bc;
            best_w[i] = bw;
            ngr[i] = ng;
            int32_t own = C[i];
            /* singleton guard, OneLevelKernel.cu:407-410 */
            if (size[own] == 1 && size[bc] == 1 && bc > own) bc = own;
            if (bc != own) {
                /* remove/insert both use W to the new community, :52-78,414-418 */
                next[i] = bc;
                atomic_add64(&ddeg[own], -L->k[i]);
                atomic_add64(&ddeg[bc], L->k[i]);
                atomic_add64(&din[own], -(bw + L->loop[i]));
                atomic_add64(&din[bc], bw + L->loop[i]);
                atomic_add32(&dsize[own], -1);
                atomic_add32(&dsize[bc], 1);
            } else {
                next[i] = own;
            }
        }
        if (pass > cap) {
            cap *= 2;
            double *nq = (double *)realloc(L->pass_q, (size_t)cap * sizeof(double));
            if (!nq) { rc = ORC_ERR_NOMEM; goto out_scratch; }
            L->pass_q = nq;
            if (R->trace) {
                int32_t **np = (int32_t **)realloc(L->pass_next, (size_t)cap * sizeof(int32_t *));
                if (!np) { rc = ORC_ERR_NOMEM; goto out_scratch; }
                L->pass_next = np;
            }
        }
        /* step 6: Q of S, before this pass's deltas, OneLevelKernel.cu:875-892 */
        double qp = q_of_state(n, in, deg, L->m, gamma);
        L->pass_q[pass - 1] = qp;
        L->passes = pass;
        int last = (qp - curQ) < R->min_gain;
        if (R->max_passes > 0 && pass >= R->max_passes) last = 1;
        if (last) {
            if (R->trace) L->pass_next[pass - 1] = NULL; /* filled after the merge */
            break;
        }
        if (R->trace) {
            L->pass_next[pass - 1] = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
            if (!L->pass_next[pass - 1]) { rc = ORC_ERR_NOMEM; goto out_scratch; }
            memcpy(L->pass_next[pass - 1], next, (size_t)n * sizeof(int32_t));
        }
        /* step 7: clamp, apply, rotate.  :894-909, :93-107 */
        curQ = qp;
        if (curQ < L->lower) curQ = L->lower;
        for (int32_t c = 0; c < n; ++c) {
            size[c] += dsize[c];
            deg[c] += ddeg[c];
            in[c] += din[c];
            dsize[c] = 0; ddeg[c] = 0; din[c] = 0;
        }
        int32_t *tmp = C; C = next; next = tmp;
    }

    /* A.5 post-loop: discard the last deltas, keep the proposals, merge
     * isolated singletons.  OneLevelKernel.cu:109-171,921-944 */
    memset(ddeg, 0, (size_t)n * sizeof(int64_t));
    memset(din, 0, (size_t)n * sizeof(int64_t));
    memset(dsize, 0, (size_t)n * sizeof(int32_t));
    for (int32_t i = 0; i < n; ++i) {
        if (ngr[i] == 0) continue;
        int32_t own = C[i];
        if (size[own] != 1) continue;
        if (ngr[i] == 1) {
            int32_t c0 = best_c[i];
            if (c0 != own && own < c0) {
                next[i] = c0;
                ddeg[own] -= L->k[i];
                ddeg[c0] += L->k[i];
                din[own] -= L->loop[i];
                din[c0] += L->loop[i];
                dsize[own] -= 1;
                dsize[c0] += 1;
            }
        }
    }
    for (int32_t c = 0; c < n; ++c) {
        size[c] += dsize[c];
        deg[c] += ddeg[c];
        in[c] += din[c];
    }
    if (R->trace) {
        L->pass_next[L->passes - 1] = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
        if (!L->pass_next[L->passes - 1]) { rc = ORC_ERR_NOMEM; goto out_scratch; }
        memcpy(L->pass_next[L->passes - 1], next, (size_t)n * sizeof(int32_t));
    }
    L->q = q_of_state(n, in, deg, L->m, gamma);
    L->c_final = next;
    next = NULL;

out_scratch:
    for (int t = 0; t < nthreads; ++t) {
        if (accs) free(accs[t]);
        if (stamps) free(stamps[t]);
        if (gbufs) free(gbufs[t]);
    }
    free(accs); free(stamps); free(gbufs); free(tokens);
out:
    free(C); free(next); free(deg); free(in); free(size);
    free(ddeg); free(din); free(dsize);
    free(best_c); free(best_w); free(ngr);
    return rc;
}

/* A.6 renumber: new id = rank of the community's smallest member position;
 * the Partition lists positions ordered by (old community, position).
 * RenumberKernel.cu:117-133,160-163,191,202-203 */
static int renumber(orc_run *R, orc_level *L, int level)
{
    int32_t n = L->n;
    int32_t *minpos = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    L->newc = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    L->part_node = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    L->part_comm = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int32_t *idmap = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    int64_t *perm = (int64_t *)xmalloc((size_t)n * sizeof(int64_t));
    if (!minpos || !L->newc || !L->part_node || !L->part_comm || !idmap || !perm) {
        free(minpos); free(idmap); free(perm);
        return ORC_ERR_NOMEM;
    }
    for (int32_t c = 0; c < n; ++c) minpos[c] = -1;
    for (int32_t i = 0; i < n; ++i)
        if (minpos[L->c_final[i]] < 0) minpos[L->c_final[i]] = i;
    int32_t K = 0;
    for (int32_t i = 0; i < n; ++i)
        if (minpos[L->c_final[i]] == i) idmap[L->c_final[i]] = K++;
    for (int32_t i = 0; i < n; ++i) L->newc[i] = idmap[L->c_final[i]];
    L->kcomm = K;
    for (int32_t i = 0; i < n; ++i) perm[i] = i;
    int rc = counting_sort_stable(n, L->c_final, n, perm);
    if (rc == ORC_OK) {
        for (int32_t j = 0; j < n; ++j) {
            int32_t p = (int32_t)perm[j];
            L->part_node[j] = (level == 0) ? R->ids[p] : p;
            L->part_comm[j] = L->newc[p];
        }
    }
    free(minpos); free(idmap); free(perm);
    return rc;
}

/* A.7 coarsen: E' = sum of w' per ordered pair (newid(C[s]), newid(C[t])),
 * lexicographically sorted.  InducedGraphKernel.cu:38-45,91-99 */
static int coarsen(const orc_level *L, orc_level *Nx)
{
    int64_t e = L->e;
    int32_t K = L->kcomm;
    int32_t *ks = (int32_t *)xmalloc((size_t)e * sizeof(int32_t));
    int32_t *kt = (int32_t *)xmalloc((size_t)e * sizeof(int32_t));
    int64_t *idx = (int64_t *)xmalloc((size_t)e * sizeof(int64_t));
    if (!ks || !kt || !idx) { free(ks); free(kt); free(idx); return ORC_ERR_NOMEM; }
    for (int64_t j = 0; j < e; ++j) {
        ks[j] = L->newc[L->es[j]];
        kt[j] = L->newc[L->et[j]];
        idx[j] = j;
    }
    int rc = counting_sort_stable(e, kt, K, idx);
    if (rc == ORC_OK) rc = counting_sort_stable(e, ks, K, idx);
    if (rc != ORC_OK) { free(ks); free(kt); free(idx); return rc; }
    int64_t cnt = 0;
    for (int64_t j = 0; j < e; ++j) {
        int64_t o = idx[j];
        if (j == 0 || ks[o] != ks[idx[j - 1]] || kt[o] != kt[idx[j - 1]]) ++cnt;
    }
    Nx->n = K;
    Nx->e = cnt;
    Nx->es = (int32_t *)xmalloc((size_t)cnt * sizeof(int32_t));
    Nx->et = (int32_t *)xmalloc((size_t)cnt * sizeof(int32_t));
    Nx->ew = (int64_t *)xmalloc((size_t)cnt * sizeof(int64_t));
    if (!Nx->es || !Nx->et || !Nx->ew) { free(ks); free(kt); free(idx); return ORC_ERR_NOMEM; }
    int64_t o2 = -1;
    for (int64_t j = 0; j < e; ++j) {
        int64_t o = idx[j];
        int64_t w = L->ew[o] == 0 ? 1 : L->ew[o];
        if (j == 0 || ks[o] != ks[idx[j - 1]] || kt[o] != kt[idx[j - 1]]) {
            ++o2;
            Nx->es[o2] = ks[o];
            Nx->et[o2] = kt[o];
            Nx->ew[o2] = 0;
        }
        Nx->ew[o2] += w;
    }
    free(ks); free(kt); free(idx);
    return ORC_OK;
}

static int grow_levels(orc_run *R)
{
    if (R->nlevels < R->cap) return ORC_OK;
    int nc = R->cap ? 2 * R->cap : 8;
    orc_level *nl = (orc_level *)realloc(R->lv, (size_t)nc * sizeof(orc_level));
    if (!nl) return ORC_ERR_NOMEM;
    memset(nl + R->cap, 0, (size_t)(nc - R->cap) * sizeof(orc_level));
    R->lv = nl;
    R->cap = nc;
    return ORC_OK;
}

static int cmp_i64pair(const void *a, const void *b)
{
    const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

/* Maps node ids to positions (LouvainCUDA/InitKernel.cu:159-169 inserts the
 * node list into a hash table keyed by id).  Unknown or repeated ids are input
 * errors here; the reference leaves them undefined. */
static int map_edges(orc_run *R, int64_t e, const int32_t *src, const int32_t *dst, orc_level *L)
{
    int32_t n = R->n0;
    int iota = 1;
    for (int32_t i = 0; i < n && iota; ++i) iota = (R->ids[i] == i);
    L->es = (int32_t *)xmalloc((size_t)e * sizeof(int32_t));
    L->et = (int32_t *)xmalloc((size_t)e * sizeof(int32_t));
    if (!L->es || !L->et) return ORC_ERR_NOMEM;
    if (iota) {
        for (int64_t j = 0; j < e; ++j) {
            if (src[j] < 0 || src[j] >= n || dst[j] < 0 || dst[j] >= n) return ORC_ERR_INPUT;
            L->es[j] = src[j];
            L->et[j] = dst[j];
        }
        return ORC_OK;
    }
    int64_t *pairs = (int64_t *)xmalloc((size_t)n * 2 * sizeof(int64_t));
    if (!pairs) return ORC_ERR_NOMEM;
    for (int32_t i = 0; i < n; ++i) { pairs[2 * i] = R->ids[i]; pairs[2 * i + 1] = i; }
    qsort(pairs, (size_t)n, 2 * sizeof(int64_t), cmp_i64pair);
    for (int32_t i = 1; i < n; ++i)
        if (pairs[2 * i] == pairs[2 * i - 2]) { free(pairs); return ORC_ERR_INPUT; }
    for (int64_t j = 0; j < e; ++j) {
        for (int side = 0; side < 2; ++side) {
            int64_t key = side ? dst[j] : src[j];
            int32_t lo = 0, hi = n - 1, found = -1;
            while (lo <= hi) {
                int32_t mid = lo + (hi - lo) / 2;
                if (pairs[2 * mid] == key) { found = (int32_t)pairs[2 * mid + 1]; break; }
                if (pairs[2 * mid] < key) lo = mid + 1; else hi = mid - 1;
            }
            if (found < 0) { free(pairs); return ORC_ERR_INPUT; }
            if (side) L->et[j] = found; else L->es[j] = found;
        }
    }
    free(pairs);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* exported API                                                              */

orc_run *orc_run_louvain(int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                         const int32_t *dst, const void *w, int wtype, double gamma,
                         double min_gain, int32_t max_passes, int trace, int threads)
{
    orc_run *R = (orc_run *)calloc(1, sizeof(orc_run));
    if (!R) return NULL;
    R->n0 = n; R->e0 = e; R->gamma = gamma; R->min_gain = min_gain;
    R->max_passes = max_passes; R->trace = trace; R->threads = threads > 0 ? threads : 1;
    if (n < 0 || e < 0 || (wtype != 0 && wtype != 1)) { R->status = ORC_ERR_ARG; return R; }
    R->ids = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    if (!R->ids) { R->status = ORC_ERR_NOMEM; return R; }
    for (int32_t i = 0; i < n; ++i) R->ids[i] = ids ? ids[i] : i;

    /* weight units */
    int s = 0;
    if (wtype == 1) {
        const float *wf = (const float *)w;
        for (int64_t j = 0; j < e; ++j) {
            int sj = float_unit_exponent(wf[j]);
            if (sj < 0) { R->status = ORC_ERR_INPUT; return R; }
            if (sj > s) s = sj;
        }
    } else {
        const int32_t *wi = (const int32_t *)w;
        for (int64_t j = 0; j < e; ++j)
            if (wi[j] < 0) { R->status = ORC_ERR_INPUT; return R; }
    }
    R->scale = s;
    progress("weights checked, e", 0, e);

    if (grow_levels(R) != ORC_OK) { R->status = ORC_ERR_NOMEM; return R; }
    orc_level *L0 = &R->lv[0];
    L0->n = n;
    L0->e = e;
    int rc = map_edges(R, e, src, dst, L0);
    if (rc != ORC_OK) { R->status = rc; return R; }
    L0->ew = (int64_t *)xmalloc((size_t)e * sizeof(int64_t));
    if (!L0->ew) { R->status = ORC_ERR_NOMEM; return R; }
    int64_t total = 0;
    for (int64_t j = 0; j < e; ++j) {
        double v = (wtype == 1) ? (double)((const float *)w)[j] : (double)((const int32_t *)w)[j];
        double u = ldexp(v, s);
        if (u >= (double)ORC_UNIT_LIMIT) { R->status = ORC_ERR_RANGE; return R; }
        L0->ew[j] = (int64_t)u;
        total += L0->ew[j] == 0 ? 1 : L0->ew[j];
        if (total >= ORC_UNIT_LIMIT) { R->status = ORC_ERR_RANGE; return R; }
    }

    R->final_comm = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    if (!R->final_comm) { R->status = ORC_ERR_NOMEM; return R; }

    if (e == 0) {
        /* D7: identity partition, one level */
        L0->kcomm = n;
        L0->c_final = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
        L0->newc = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
        L0->part_node = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
        L0->part_comm = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
        if (!L0->c_final || !L0->newc || !L0->part_node || !L0->part_comm) { R->status = ORC_ERR_NOMEM; return R; }
        for (int32_t i = 0; i < n; ++i) {
            L0->c_final[i] = i; L0->newc[i] = i;
            L0->part_node[i] = R->ids[i]; L0->part_comm[i] = i;
            R->final_comm[i] = i;
        }
        R->nlevels = 1;
        R->final_q = 0.0;
        return R;
    }

    /* A.8 level loop, LouvainCPP.cpp:334-507 */
    double mod = 0.0;
    int level = 0;
    for (;;) {
        orc_level *L = &R->lv[level];
        progress("csr", level, L->e);
        rc = build_csr(L);
        if (rc == ORC_OK) rc = init_vertices(L);
        if (rc != ORC_OK) { R->status = rc; return R; }
        progress("init done, n", level, L->n);
        {
            /* initial Q of this level: internals = loops, degrees = k */
            L->q_init = q_of_state(L->n, L->loop, L->k, L->m, gamma);
        }
        if (level == 0) mod = L->q_init;   /* LouvainCPP.cpp:356 */
        L->lower = mod;                     /* :366 and :446 */
        rc = one_level(R, L);
        if (rc == ORC_OK) rc = renumber(R, L, level);
        if (rc != ORC_OK) { R->status = rc; return R; }
        progress("renumbered, k", level, L->kcomm);
        R->nlevels = level + 1;             /* push, :384 / :464 */
        double q = L->q;
        if (level > 0) {
            if (q - mod < min_gain) break;  /* :466-471 */
        }
        mod = q;                            /* :386 / :473 */
        if (grow_levels(R) != ORC_OK) { R->status = ORC_ERR_NOMEM; return R; }
        L = &R->lv[level];
        rc = coarsen(L, &R->lv[level + 1]);
        if (rc != ORC_OK) { R->status = rc; return R; }
        progress("coarsened, e'", level, R->lv[level + 1].e);
        ++level;
    }

    /* flatten at the last level */
    for (int32_t i = 0; i < n; ++i) {
        int32_t f = R->lv[0].newc[i];
        for (int l = 1; l < R->nlevels; ++l) f = R->lv[l].newc[f];
        R->final_comm[i] = f;
    }
    return R;
}

int orc_status(const orc_run *R) { return R ? R->status : ORC_ERR_NOMEM; }
int orc_scale(const orc_run *R) { return R->scale; }
int orc_num_levels(const orc_run *R) { return R->nlevels; }

static const orc_level *lvl(const orc_run *R, int l)
{
    if (!R || l < 0 || l >= R->nlevels) return NULL;
    return &R->lv[l];
}

int32_t orc_level_n(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->n : -1; }
int64_t orc_level_e(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->e : -1; }
int64_t orc_level_m(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->m : -1; }
int32_t orc_level_k(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->kcomm : -1; }
int32_t orc_level_passes(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->passes : -1; }
double orc_level_q(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->q : NAN; }
double orc_level_q_init(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->q_init : NAN; }
double orc_level_lower(const orc_run *R, int l) { const orc_level *L = lvl(R, l); return L ? L->lower : NAN; }

double orc_level_pass_q(const orc_run *R, int l, int p)
{
    const orc_level *L = lvl(R, l);
    if (!L || p < 1 || p > L->passes) return NAN;
    return L->pass_q[p - 1];
}

int orc_level_pass_next(const orc_run *R, int l, int p, int32_t *out)
{
    const orc_level *L = lvl(R, l);
    if (!L || !L->pass_next || p < 1 || p > L->passes || !L->pass_next[p - 1]) return ORC_ERR_ARG;
    memcpy(out, L->pass_next[p - 1], (size_t)L->n * sizeof(int32_t));
    return ORC_OK;
}

int orc_level_newc(const orc_run *R, int l, int32_t *out)
{
    const orc_level *L = lvl(R, l);
    if (!L) return ORC_ERR_ARG;
    memcpy(out, L->newc, (size_t)L->n * sizeof(int32_t));
    return ORC_OK;
}

int orc_level_cfinal(const orc_run *R, int l, int32_t *out)
{
    const orc_level *L = lvl(R, l);
    if (!L) return ORC_ERR_ARG;
    memcpy(out, L->c_final, (size_t)L->n * sizeof(int32_t));
    return ORC_OK;
}

int orc_level_partition(const orc_run *R, int l, int32_t *node, int32_t *comm)
{
    const orc_level *L = lvl(R, l);
    if (!L) return ORC_ERR_ARG;
    memcpy(node, L->part_node, (size_t)L->n * sizeof(int32_t));
    memcpy(comm, L->part_comm, (size_t)L->n * sizeof(int32_t));
    return ORC_OK;
}

int orc_level_edges(const orc_run *R, int l, int32_t *s, int32_t *t, int64_t *w)
{
    const orc_level *L = lvl(R, l);
    if (!L) return ORC_ERR_ARG;
    memcpy(s, L->es, (size_t)L->e * sizeof(int32_t));
    memcpy(t, L->et, (size_t)L->e * sizeof(int32_t));
    memcpy(w, L->ew, (size_t)L->e * sizeof(int64_t));
    return ORC_OK;
}

/* the reference CSR of the level: rowptr[n+1], 2E (col, w) incl. self entries */
int orc_level_csr(const orc_run *R, int l, int64_t *rowptr, int32_t *col, int64_t *w)
{
    const orc_level *L = lvl(R, l);
    if (!L || !L->rowptr) return ORC_ERR_ARG;
    memcpy(rowptr, L->rowptr, (size_t)(L->n + 1) * sizeof(int64_t));
    memcpy(col, L->col, (size_t)(2 * L->e) * sizeof(int32_t));
    memcpy(w, L->cw, (size_t)(2 * L->e) * sizeof(int64_t));
    return ORC_OK;
}

int orc_level_vertex(const orc_run *R, int l, int64_t *k, int64_t *loop)
{
    const orc_level *L = lvl(R, l);
    if (!L || !L->k) return ORC_ERR_ARG;
    memcpy(k, L->k, (size_t)L->n * sizeof(int64_t));
    memcpy(loop, L->loop, (size_t)L->n * sizeof(int64_t));
    return ORC_OK;
}

/* partitionAtLevel (LouvainCPP.cpp:82-106): level-0 Partition order, each
 * community mapped through levels 1..level.  Unlike the reference it does not
 * write through the level-0 arrays. */
int orc_flatten(const orc_run *R, int level, int32_t *node, int32_t *comm)
{
    if (!R || level < 0 || level >= R->nlevels) return ORC_ERR_ARG;
    const orc_level *L0 = &R->lv[0];
    for (int32_t j = 0; j < L0->n; ++j) {
        int32_t c = L0->part_comm[j];
        for (int l = 1; l <= level; ++l) c = R->lv[l].newc[c];
        node[j] = L0->part_node[j];
        comm[j] = c;
    }
    return ORC_OK;
}

int orc_final_comm(const orc_run *R, int32_t *out)
{
    if (!R || R->status != ORC_OK) return ORC_ERR_ARG;
    memcpy(out, R->final_comm, (size_t)R->n0 * sizeof(int32_t));
    return ORC_OK;
}

/* Newman-Girvan modularity of a partition on the input graph, plain FP64:
 * Q = sum_c L_c/m - gamma*(d_c/(2m))^2, a self-loop adding w to L_c and 2w to
 * d_c (the convention of networkx.algorithms.community.modularity).
 * comm[] is indexed by position (the order of ids). */
double orc_modularity_fp64(int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                           const int32_t *dst, const void *w, int wtype, const int32_t *comm,
                           double gamma)
{
    orc_run tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.n0 = n;
    tmp.ids = (int32_t *)xmalloc((size_t)n * sizeof(int32_t));
    if (!tmp.ids) return NAN;
    for (int32_t i = 0; i < n; ++i) tmp.ids[i] = ids ? ids[i] : i;
    orc_level L;
    memset(&L, 0, sizeof(L));
    if (map_edges(&tmp, e, src, dst, &L) != ORC_OK) { free(tmp.ids); free(L.es); free(L.et); return NAN; }
    double *lc = (double *)xcalloc((size_t)n, sizeof(double));
    double *dc = (double *)xcalloc((size_t)n, sizeof(double));
    double m = 0.0;
    for (int64_t j = 0; j < e; ++j) {
        double v = (wtype == 1) ? (double)((const float *)w)[j] : (double)((const int32_t *)w)[j];
        int32_t a = comm[L.es[j]], b = comm[L.et[j]];
        m += v;
        dc[a] += v;
        dc[b] += v;
        if (a == b) lc[a] += v;
    }
    double q = 0.0;
    if (m > 0) {
        for (int32_t c = 0; c < n; ++c) {
            double x = dc[c] / (2.0 * m);
            q += lc[c] / m - gamma * x * x;
        }
    }
    free(lc); free(dc); free(tmp.ids); free(L.es); free(L.et);
    return q;
}

double orc_final_modularity(const orc_run *R, const int32_t *src, const int32_t *dst,
                            const void *w, int wtype)
{
    if (!R || R->status != ORC_OK) return NAN;
    return orc_modularity_fp64(R->n0, R->ids, R->e0, src, dst, w, wtype, R->final_comm, R->gamma);
}

void orc_free(orc_run *R)
{
    if (!R) return;
    for (int l = 0; l < R->cap; ++l) level_free(&R->lv[l], R->trace);
    free(R->lv);
    free(R->ids);
    free(R->final_comm);
    free(R);
}
