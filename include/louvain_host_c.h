/*
 * louvain_host_c.h - flat C entry points of liblouvain_host.so, for callers
 * that reach the host layer through an FFI (ctypes, cgo, JNI) instead of C++.
 *
 * lvh_*  wrap generate_dendogram and the dendrogram queries of the reference
 *        (LouvainCUDA/LouvainCPP.cpp:82-127,209-553); they replace the
 *        Boost.Python module (LouvainCUDA/LouvainCPP.cpp:943-951) as the
 *        language boundary.
 * lvg_*  deterministic synthetic graphs (the reference ships none): SBM and
 *        R-MAT edge lists, canonical (u <= v) and lexicographically sorted.
 * lvh_load_text / lvh_binary_*  graph files: the reference CLI's nodes.txt and
 *        edges.txt (LouvainCUDA/LouvainCPP.cpp:660-737) and a binary edge list.
 */
#ifndef LOUVAIN_HOST_C_H
#define LOUVAIN_HOST_C_H

#include <stdint.h>

#include "louvain_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lvh_dendrogram lvh_dendrogram;
typedef struct lvg_graph lvg_graph;

/* generate_dendogram over an edge list with float weights (ids NULL = 0..n-1). */
int lvh_generate_dendrogram(int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                            const int32_t *dst, const float *w, const lv_config *cfg,
                            lvh_dendrogram **out);
int lvh_num_levels(const lvh_dendrogram *d);
int32_t lvh_level_count(const lvh_dendrogram *d, int level);
/* Partition of one level, reference layout (node, community). */
int lvh_level_partition(const lvh_dendrogram *d, int level, int32_t *node, int32_t *comm);
/* partitionAtLevel (LouvainCUDA/LouvainCPP.cpp:82-106). */
int lvh_partition_at_level(lvh_dendrogram *d, int level, int32_t *node, int32_t *comm);
/* getPartitions (LouvainCUDA/LouvainCPP.cpp:108-114). */
int lvh_get_partitions(lvh_dendrogram *d, int level, int32_t *node, int32_t *comm);
/* countPartitions of partitionAtLevel(level) (LouvainCUDA/LouvainCPP.cpp:116-127). */
int lvh_count_partitions(lvh_dendrogram *d, int level);
void lvh_free(lvh_dendrogram *d);
/* calculateClusterSize + layoutProcess (LouvainCUDA/LouvainCPP.cpp:555-651) for the
 * graph with node ids ids[0..n-1] and a Partition (node, comm) of count entries:
 * cluster_size[nclusters] (inclusive prefix sums), and per node (ids order) its
 * cluster, index in cluster and circle position. */
int lvh_cluster_layout(int32_t n, const int32_t *ids, int32_t count, const int32_t *node,
                       const int32_t *comm, int32_t nclusters, int32_t *cluster_size,
                       int32_t *cluster, int32_t *index_in_cluster, float *x, float *y);

lvg_graph *lvg_sbm(int32_t n, int32_t block_size, double p_in, double p_out, uint64_t seed);
lvg_graph *lvg_rmat(int32_t scale, int64_t edge_factor, double a, double b, double c,
                    uint64_t seed, int wtype);
int32_t lvg_num_nodes(const lvg_graph *g);
int64_t lvg_num_edges(const lvg_graph *g);
int lvg_wtype(const lvg_graph *g);
int lvg_copy(const lvg_graph *g, int32_t *src, int32_t *dst, void *w);
void lvg_free(lvg_graph *g);

/* loadNodes + loadEdges (LouvainCUDA/LouvainCPP.cpp:660-737), the rules of
 * loadNodesFile/loadEdgesFile in louvain_host.hpp: node ids and float
 * weights; edges name node ids.  LV_ERR_ARG when a file cannot be opened,
 * LV_ERR_INVALID_INPUT when its content breaks the rules. */
typedef struct lvh_file_graph lvh_file_graph;
int lvh_load_text(const char *nodes_path, const char *edges_path, lvh_file_graph **out);
int32_t lvh_file_num_nodes(const lvh_file_graph *g);
int64_t lvh_file_num_edges(const lvh_file_graph *g);
/* ids[num_nodes], src/dst/w[num_edges] (ids may be NULL). */
int lvh_file_copy(const lvh_file_graph *g, int32_t *ids, int32_t *src, int32_t *dst, float *w);
void lvh_file_free(lvh_file_graph *g);

/* Binary edge list, little-endian: a 32-byte header - "LVEB", u32 version 1,
 * i32 n, u32 flags (bit 0: node ids present, bit 1: float32 weights, else
 * int32), i64 e, 8 zero bytes - then ids int32[n] (if flagged), src int32[e],
 * dst int32[e], w int32/float32[e].  No text parse: for graphs of billions of
 * edges.  lvh_binary_header checks the header against the file size;
 * lvh_binary_read fills the caller's arrays (ids may be NULL); lvh_binary_write
 * writes one (ids NULL: none stored). */
int lvh_binary_header(const char *path, int32_t *n, int64_t *e, int *wtype, int *has_ids);
int lvh_binary_read(const char *path, int32_t *ids, int32_t *src, int32_t *dst, void *w);
int lvh_binary_write(const char *path, int32_t n, const int32_t *ids, int64_t e,
                     const int32_t *src, const int32_t *dst, const void *w, int wtype);

#ifdef __cplusplus
}
#endif

#endif /* LOUVAIN_HOST_C_H */
