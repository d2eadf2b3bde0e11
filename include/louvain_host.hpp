/*
 * louvain_host.hpp - C++ host API of liblouvain_host.so.
 *
 * Mirrors the reference's host types and level loop so that a caller of the
 * reference can switch over unchanged:
 *   Edge, EdgeContainer   LouvainCUDA/Edge.h:26-69
 *   Graph                 LouvainCUDA/Graph.h:39-130 (storage and accessors; the
 *                         visualisation fields x/y/cluster are not part of the path)
 *   Community, Partition  LouvainCUDA/community.h:16-20, LouvainCUDA/partition.h:14-29
 *   generate_dendogram    LouvainCUDA/LouvainCPP.cpp:209-553
 *   partitionAtLevel, getPartitions, countPartitions
 *                         LouvainCUDA/LouvainCPP.cpp:82-127
 * generate_dendogram drives the device engine only through the C ABI in
 * louvain_amd.h, phase by phase, in the reference's call order.
 *
 * Errors: the reference reports none; here a failing engine call throws
 * louvain::Error (a std::runtime_error carrying the lv_status code), and
 * getPartitions keeps the reference's `throw const char*` for a bad level.
 */
#ifndef LOUVAIN_HOST_HPP
#define LOUVAIN_HOST_HPP

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "louvain_amd.h"

struct Edge {
    int target, source;
    float weight;
};

class EdgeContainer {
public:
    EdgeContainer() {}
    ~EdgeContainer();
    EdgeContainer(const EdgeContainer &) = delete;
    EdgeContainer &operator=(const EdgeContainer &) = delete;
    void allocate(int count);
    int *target = nullptr;
    int *source = nullptr;
    float *weight = nullptr;
    int count = 0;
};

class Graph {
public:
    Graph();
    ~Graph();
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;

    void allocateNodes(int size);
    void allocateEdges(int size);
    void addNode(int node);
    void addEdge(Edge edge);
    void addEdgeFromVector(std::vector<Edge> edges);
    int setNodes(int *nodes, int size);
    int setEdges(Edge *edges, int size);

    Edge getEdgeAt(int idx);
    int getNodeAt(int idx);
    int *getNodes();
    EdgeContainer &getEdgeContainer();
    int getNodeSize();
    int getEdgeSize();
    int getIdxOfNode(int node);

    /* Per-node cluster, index within the cluster and circle position, filled by
     * calculateClusterSize and layoutProcess (the reference's Graph.h keeps
     * them per node too). */
    void setCluster(int idx, int cluster);
    int getCluster(int idx);
    void setIndexInCluster(int idx, int index);
    int getIndexInCluster(int idx);
    float getX(int idx);
    float getY(int idx);
    float *getXPtr(int idx);
    float *getYPtr(int idx);

private:
    void ensureNodeAttrs();
    int *nodes = nullptr;
    int nodeCapacity = 0;
    int nodeSize = 0;
    int edgeSize = 0;
    EdgeContainer edges;
    std::map<int, int> nodeMap;
    std::vector<int> cluster_, indexInCluster_;
    std::vector<float> x_, y_;
};

struct Community {
    int *node = nullptr;
    int *community = nullptr;
};

struct Partition {
    int count = 0;
    Community partition;
    void Deallocate();
};

namespace louvain {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &where);
    int status() const { return status_; }

private:
    int status_;
};

}  // namespace louvain

/* generate_dendogram (LouvainCUDA/LouvainCPP.cpp:209-553).  part_init is accepted
 * and, as in the reference, not used.  One Partition per level; the caller
 * frees them with Partition::Deallocate.  The overload with an lv_config sets
 * resolution, thresholds and the device. */
std::vector<Partition> generate_dendogram(Graph &original_graph,
                                          std::map<int, std::pair<int, float>> &part_init);
std::vector<Partition> generate_dendogram(Graph &original_graph,
                                          std::map<int, std::pair<int, float>> &part_init,
                                          const lv_config &cfg);

/* partitionAtLevel (LouvainCPP.cpp:82-106): level-0 order, communities mapped
 * through levels 1..level.  Returns new arrays (the reference writes through
 * the level-0 arrays). */
Partition partitionAtLevel(std::vector<Partition> &dendogram, int level);

/* getPartitions (LouvainCPP.cpp:108-114), including its `level || size-1`
 * expression: the flattening is taken at level 1 whenever the dendrogram has
 * two or more levels. */
Partition getPartitions(std::vector<Partition> &dendogram, int level = 0);

/* The flattening at the last level (the partition Louvain converged to). */
Partition getFinalPartition(std::vector<Partition> &dendogram);

/* countPartitions (LouvainCPP.cpp:116-127): 1 + the largest community id. */
int countPartitions(Partition partitions);

/* calculateClusterSize (LouvainCPP.cpp:555-583): each node's cluster (from the
 * partition, by node id) and index among earlier nodes of its cluster; returns
 * new int[numberOfClusters] with the inclusive prefix sums of the cluster sizes
 * (delete[] by the caller).  A node missing from the partition throws
 * louvain::Error (the reference reads an uninitialised cluster). */
int *calculateClusterSize(Graph &graph, Partition partitions, int numberOfClusters);

/* layoutProcess (LouvainCPP.cpp:585-651): node i of cluster c goes to slot
 * clusterSize[c-1] + indexInCluster(i) of a circle of radius 50 centred at
 * (0, 50): x = 50 cos(2 pi slot / n), y = 50 + 50 sin(2 pi slot / n). */
void layoutProcess(int *clusterSize, Graph &graph, int nrPartitions);

/* Text loaders of the reference CLI (LouvainCPP.cpp:660-737): a count, then the
 * ids (nodes.txt) or `source target weight` triples (edges.txt), separated by
 * any whitespace, parsed with atoi/atof.  As in the reference, a node id that
 * ends the file (no trailing whitespace) is not read.  Unlike it, more entries
 * than the count (an overrun there) and an incomplete last triple (filled from
 * a stale token there) throw std::runtime_error, as does a missing file. */
std::vector<int> loadNodesFile(const std::string &path);
std::vector<Edge> loadEdgesFile(const std::string &path);

#endif /* LOUVAIN_HOST_HPP */
