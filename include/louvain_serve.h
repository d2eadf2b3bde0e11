/*
 * louvain_serve.h - a serving loop over the graph-set engine (liblouvain_host.so).
 *
 * The reference runs one graph per call and has no server (its Python entry
 * point start(), LouvainCUDA/LouvainCPP.cpp:808-941, is synchronous).  This is
 * continuous batching: requests (one graph at one resolution each) queue on
 * the host; each of W workers owns one engine context (its own HIP stream and
 * HBM arena) and repeatedly takes the waiting requests that fit its HBM budget
 * - an estimate of each graph's footprint, admitted in arrival order - runs
 * them as one graph set (lv_batch_run: CSR, community state and dendrogram
 * stay on the device across levels), and publishes the results.  Requests
 * arriving while a set runs are admitted by the next free worker, so sets
 * overlap on the device.  Components of a set are independent, so every
 * result equals the request's result alone.
 *
 * Times are milliseconds of a steady clock since the server was created.
 */
#ifndef LOUVAIN_SERVE_H
#define LOUVAIN_SERVE_H

#include <stdint.h>

#include "louvain_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lvs_server lvs_server;

typedef struct {
    int32_t status;       /* LV_OK or the engine status of the request's own run */
    int32_t levels;       /* dendrogram levels */
    int32_t communities;  /* communities of the final partition */
    int32_t batch_size;   /* requests in the set it ran in */
    double q;             /* modularity returned by the last level */
    double t_submit_ms, t_start_ms, t_done_ms;
} lvs_result;

typedef struct {
    int64_t submitted, completed, sets;
    int32_t max_set;      /* most requests in one set */
    int32_t workers;
} lvs_stats;

/* workers contexts on cfg->device, each with an arena of budget_bytes (0: 32 GiB).
 * cfg->gamma is ignored (each request carries its resolution). */
int lvs_create(lvs_server **out, const lv_config *cfg, int32_t workers, uint64_t budget_bytes);
/* Queues the graph (ids 0..n-1); *id identifies the request.  The arrays are
 * not copied: they must stay valid and unchanged until lvs_wait returns for
 * this request (the worker uploads them straight to the device).  A graph whose
 * footprint estimate (200 bytes per edge + 320 per vertex) exceeds the budget
 * is refused here with LV_ERR_NOMEM.  Invalid input (an id outside 0..n-1, a
 * negative or non-finite weight) is reported by lvs_wait as that request's
 * status; it does not affect the other requests of its set. */
int lvs_submit(lvs_server *s, int32_t n, int64_t e, const int32_t *src, const int32_t *dst,
               const void *w, int wtype, double gamma, int64_t *id);
/* Blocks until the request is done; copies its final partition (n entries,
 * comm may be NULL) and summary, then forgets the request.  Returns the
 * request's status; LV_ERR_ARG for an unknown id or one that another lvs_wait
 * is already waiting for. */
int lvs_wait(lvs_server *s, int64_t id, int32_t *comm, lvs_result *res);
int lvs_get_stats(lvs_server *s, lvs_stats *out);
/* Milliseconds since the server was created (the clock of lvs_result). */
double lvs_now_ms(lvs_server *s);
/* Finishes the queued requests, stops the workers, frees everything. */
int lvs_destroy(lvs_server *s);

#ifdef __cplusplus
}
#endif

#endif /* LOUVAIN_SERVE_H */
