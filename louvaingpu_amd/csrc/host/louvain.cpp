// louvain.cpp - generate_dendogram and the dendrogram queries of the reference
// (LouvainCUDA/LouvainCPP.cpp:82-127,209-553), over the engine's C ABI; plus a
// flat C API (lvh_*) for callers that cannot use C++ (ctypes, FFI).
#include <cstring>
#include <memory>
#include <string>
#include <cmath>
#include <unordered_map>

#include "louvain_host.hpp"

void Partition::Deallocate()
{
    delete[] partition.node;
    delete[] partition.community;
    partition.node = nullptr;
    partition.community = nullptr;
}

namespace louvain {

Error::Error(int status, const std::string &where)
    : std::runtime_error(where + ": " + lv_strerror(status)), status_(status)
{
}

}  // namespace louvain

namespace {

#define LVH_CALL(expr)                                                   \
    do {                                                                 \
        int lvh_rc_ = (expr);                                            \
        if (lvh_rc_ != LV_OK) throw louvain::Error(lvh_rc_, #expr);      \
    } while (0)

struct CtxGuard {
    lv_ctx *c = nullptr;
    ~CtxGuard()
    {
        if (c) lv_destroy(c);
    }
};

Partition fetch_partition(lv_ctx *c, int32_t level)
{
    lv_level_info info;
    LVH_CALL(lv_level_info_get(c, level, &info));
    Partition p;
    p.count = info.n;
    p.partition.node = new int[info.n > 0 ? info.n : 1];
    p.partition.community = new int[info.n > 0 ? info.n : 1];
    int rc = lv_level_partition(c, level, p.partition.node, p.partition.community);
    if (rc != LV_OK) {
        p.Deallocate();
        throw louvain::Error(rc, "lv_level_partition");
    }
    return p;
}

// Identity partition for an edgeless graph (python-louvain's convention; the
// reference leaves the arrays uninitialised, LouvainCPP.cpp:211-230).
std::vector<Partition> identity_dendrogram(Graph &g)
{
    Partition p;
    p.count = g.getNodeSize();
    p.partition.node = new int[p.count > 0 ? p.count : 1];
    p.partition.community = new int[p.count > 0 ? p.count : 1];
    for (int i = 0; i < p.count; ++i) {
        p.partition.node[i] = g.getNodeAt(i);
        p.partition.community[i] = i;
    }
    return std::vector<Partition>{p};
}

}  // namespace

std::vector<Partition> generate_dendogram(Graph &original_graph,
                                          std::map<int, std::pair<int, float>> &part_init,
                                          const lv_config &cfg)
{
    (void)part_init;  // accepted and unused, as in the reference (LouvainCPP.cpp:209)
    if (original_graph.getEdgeSize() == 0) return identity_dendrogram(original_graph);
    std::vector<Partition> status_list;
    CtxGuard g;
    LVH_CALL(lv_create(&g.c, &cfg));
    lv_ctx *c = g.c;
    EdgeContainer &ec = original_graph.getEdgeContainer();
    LVH_CALL(lv_load_edges(c, original_graph.getNodeSize(), original_graph.getNodes(),
                           original_graph.getEdgeSize(), ec.source, ec.target, ec.weight,
                           LV_WEIGHT_FLOAT32));
    try {
        // level 0 (LouvainCPP.cpp:334-401)
        LVH_CALL(lv_neighbour_compute(c));
        LVH_CALL(lv_init(c));
        double mod = 0.0, new_mod = 0.0;
        LVH_CALL(lv_modularity(c, &mod));
        LVH_CALL(lv_one_level(c, mod, &new_mod, nullptr));
        LVH_CALL(lv_renumber(c, nullptr));
        status_list.push_back(fetch_partition(c, 0));
        mod = new_mod;
        LVH_CALL(lv_induce(c));
        // later levels (LouvainCPP.cpp:420-507)
        for (int32_t level = 1;; ++level) {
            LVH_CALL(lv_neighbour_compute(c));
            LVH_CALL(lv_init(c));
            LVH_CALL(lv_one_level(c, mod, &new_mod, nullptr));
            LVH_CALL(lv_renumber(c, nullptr));
            status_list.push_back(fetch_partition(c, level));
            if (new_mod - mod < cfg.min_gain) break;
            if (cfg.max_levels > 0 && level + 1 >= cfg.max_levels) break;
            mod = new_mod;
            LVH_CALL(lv_induce(c));
        }
    } catch (...) {
        for (auto &p : status_list) p.Deallocate();
        throw;
    }
    return status_list;
}

std::vector<Partition> generate_dendogram(Graph &original_graph,
                                          std::map<int, std::pair<int, float>> &part_init)
{
    lv_config cfg;
    lv_config_default(&cfg);
    return generate_dendogram(original_graph, part_init, cfg);
}

Partition partitionAtLevel(std::vector<Partition> &dendogram, int level)
{
    if (dendogram.empty() || level < 0 || level >= (int)dendogram.size())
        throw std::out_of_range("partitionAtLevel: level out of range");
    const Partition &p0 = dendogram[0];
    Partition out;
    out.count = p0.count;
    out.partition.node = new int[p0.count > 0 ? p0.count : 1];
    out.partition.community = new int[p0.count > 0 ? p0.count : 1];
    std::memcpy(out.partition.node, p0.partition.node, sizeof(int) * (size_t)p0.count);
    std::memcpy(out.partition.community, p0.partition.community, sizeof(int) * (size_t)p0.count);
    for (int l = 1; l <= level; ++l) {
        // level l lists every node 0..count-1 once: invert it into a lookup
        const Partition &pl = dendogram[l];
        std::vector<int> map((size_t)pl.count, -1);
        for (int k = 0; k < pl.count; ++k)
            if (pl.partition.node[k] >= 0 && pl.partition.node[k] < pl.count)
                map[(size_t)pl.partition.node[k]] = pl.partition.community[k];
        for (int j = 0; j < out.count; ++j) {
            int com = out.partition.community[j];
            out.partition.community[j] = (com >= 0 && com < pl.count) ? map[(size_t)com] : -1;
        }
    }
    return out;
}

Partition getPartitions(std::vector<Partition> &dendogram, int level)
{
    if ((level < 0) || (level > (int)dendogram.size() - 1))
        throw "Invalid argument: level is not between 1 and len(dendogram) - 1 included.";
    return partitionAtLevel(dendogram, (level || dendogram.size() - 1) ? 1 : 0);
}

Partition getFinalPartition(std::vector<Partition> &dendogram)
{
    return partitionAtLevel(dendogram, (int)dendogram.size() - 1);
}

int countPartitions(Partition partitions)
{
    int mx = 0;
    for (int i = 0; i < partitions.count; ++i)
        if (partitions.partition.community[i] > mx) mx = partitions.partition.community[i];
    return 1 + mx;
}

int *calculateClusterSize(Graph &graph, Partition partitions, int numberOfClusters)
{
    if (numberOfClusters < 0) throw louvain::Error(LV_ERR_ARG, "calculateClusterSize");
    std::unordered_map<int, int> comm;
    comm.reserve((size_t)partitions.count * 2);
    for (int j = 0; j < partitions.count; ++j)  // the first entry of a node wins, as the
        comm.emplace(partitions.partition.node[j], partitions.partition.community[j]);  // scan does
    int *clusterSize = new int[numberOfClusters > 0 ? numberOfClusters : 1]();
    const int n = graph.getNodeSize();
    for (int i = 0; i < n; ++i) {
        auto it = comm.find(graph.getNodeAt(i));
        if (it == comm.end() || it->second < 0 || it->second >= numberOfClusters) {
            delete[] clusterSize;
            throw louvain::Error(LV_ERR_INVALID_INPUT, "calculateClusterSize");
        }
        const int c = it->second;
        graph.setCluster(i, c);
        graph.setIndexInCluster(i, clusterSize[c]);
        clusterSize[c] += 1;
    }
    for (int i = 1; i < numberOfClusters; ++i) clusterSize[i] += clusterSize[i - 1];
    return clusterSize;
}

void layoutProcess(int *clusterSize, Graph &graph, int nrPartitions)
{
    const int n = graph.getNodeSize();
    if (n > 0) graph.getXPtr(0);  // sizes the per-node arrays before the threads start
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int c = graph.getCluster(i);
        if (c < 0 || c >= nrPartitions) continue;
        const int slot = (c > 0 ? clusterSize[c - 1] : 0) + graph.getIndexInCluster(i);
        *graph.getXPtr(i) = (float)(50 * cos(2 * M_PI * slot / n));
        *graph.getYPtr(i) = (float)(50 + 50 * sin(2 * M_PI * slot / n));
    }
}

// ---------------------------------------------------------------------------
// flat C API

struct lvh_dendrogram {
    std::vector<Partition> levels;
    ~lvh_dendrogram()
    {
        for (auto &p : levels) p.Deallocate();
    }
};

extern "C" {

// Runs generate_dendogram on an edge list with float weights.  ids may be NULL
// (node ids 0..n-1).  On success *out owns the dendrogram (free with lvh_free).
int lvh_generate_dendrogram(int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                            const int32_t *dst, const float *w, const lv_config *cfg,
                            lvh_dendrogram **out)
{
    if (!out || n < 0 || e < 0 || e > INT32_MAX) return LV_ERR_ARG;
    *out = nullptr;
    try {
        Graph g;
        std::vector<int> nodes((size_t)n);
        for (int32_t i = 0; i < n; ++i) nodes[(size_t)i] = ids ? ids[i] : i;
        g.setNodes(nodes.data(), n);
        g.allocateEdges((int)e);
        for (int64_t j = 0; j < e; ++j) g.addEdge(Edge{dst[j], src[j], w[j]});
        std::map<int, std::pair<int, float>> init;
        lv_config c;
        if (cfg)
            c = *cfg;
        else
            lv_config_default(&c);
        std::unique_ptr<lvh_dendrogram> d(new lvh_dendrogram());
        d->levels = generate_dendogram(g, init, c);
        *out = d.release();
        return LV_OK;
    } catch (const louvain::Error &err) {
        return err.status();
    } catch (const std::bad_alloc &) {
        return LV_ERR_NOMEM;
    } catch (...) {
        return LV_ERR_ARG;
    }
}

int lvh_num_levels(const lvh_dendrogram *d) { return d ? (int)d->levels.size() : -1; }

int32_t lvh_level_count(const lvh_dendrogram *d, int level)
{
    if (!d || level < 0 || level >= (int)d->levels.size()) return -1;
    return d->levels[(size_t)level].count;
}

int lvh_level_partition(const lvh_dendrogram *d, int level, int32_t *node, int32_t *comm)
{
    if (!d || level < 0 || level >= (int)d->levels.size() || !node || !comm) return LV_ERR_ARG;
    const Partition &p = d->levels[(size_t)level];
    std::memcpy(node, p.partition.node, sizeof(int32_t) * (size_t)p.count);
    std::memcpy(comm, p.partition.community, sizeof(int32_t) * (size_t)p.count);
    return LV_OK;
}

static int copy_out(Partition p, int32_t *node, int32_t *comm)
{
    std::memcpy(node, p.partition.node, sizeof(int32_t) * (size_t)p.count);
    std::memcpy(comm, p.partition.community, sizeof(int32_t) * (size_t)p.count);
    p.Deallocate();
    return LV_OK;
}

int lvh_partition_at_level(lvh_dendrogram *d, int level, int32_t *node, int32_t *comm)
{
    if (!d || !node || !comm || level < 0 || level >= (int)d->levels.size()) return LV_ERR_ARG;
    return copy_out(partitionAtLevel(d->levels, level), node, comm);
}

int lvh_get_partitions(lvh_dendrogram *d, int level, int32_t *node, int32_t *comm)
{
    if (!d || !node || !comm || level < 0 || level >= (int)d->levels.size()) return LV_ERR_ARG;
    return copy_out(getPartitions(d->levels, level), node, comm);
}

int lvh_count_partitions(lvh_dendrogram *d, int level)
{
    if (!d || level < 0 || level >= (int)d->levels.size()) return LV_ERR_ARG;
    Partition p = partitionAtLevel(d->levels, level);
    int k = countPartitions(p);
    p.Deallocate();
    return k;
}

void lvh_free(lvh_dendrogram *d) { delete d; }

int lvh_cluster_layout(int32_t n, const int32_t *ids, int32_t count, const int32_t *node,
                       const int32_t *comm, int32_t nclusters, int32_t *cluster_size,
                       int32_t *cluster, int32_t *index_in_cluster, float *x, float *y)
{
    if (n < 0 || count < 0 || nclusters < 0 || (n > 0 && (!ids || !cluster || !index_in_cluster ||
                                                           !x || !y)) ||
        (count > 0 && (!node || !comm)) || (nclusters > 0 && !cluster_size))
        return LV_ERR_ARG;
    try {
        Graph g;
        g.setNodes(const_cast<int *>(ids), n);
        Partition p;
        p.count = count;
        p.partition.node = const_cast<int *>(node);
        p.partition.community = const_cast<int *>(comm);
        int *cs = calculateClusterSize(g, p, nclusters);
        layoutProcess(cs, g, nclusters);
        for (int32_t k = 0; k < nclusters; ++k) cluster_size[k] = cs[k];
        delete[] cs;
        for (int32_t i = 0; i < n; ++i) {
            cluster[i] = g.getCluster(i);
            index_in_cluster[i] = g.getIndexInCluster(i);
            x[i] = g.getX(i);
            y[i] = g.getY(i);
        }
    } catch (const louvain::Error &e) {
        return e.status();
    } catch (...) {
        return LV_ERR_NOMEM;
    }
    return LV_OK;
}

}  // extern "C"
