// graph.cpp - host Graph/EdgeContainer storage (LouvainCUDA/Graph.cpp:28-160,
// 218-347, LouvainCUDA/Edge.h:26-63) and the CLI text loaders
// (LouvainCUDA/LouvainCPP.cpp:660-737).
#include <cstdlib>
#include <fstream>
#include <stdexcept>

#include "louvain_host.hpp"

void EdgeContainer::allocate(int n)
{
    delete[] target;
    delete[] source;
    delete[] weight;
    count = n;
    target = new int[n > 0 ? n : 1];
    source = new int[n > 0 ? n : 1];
    weight = new float[n > 0 ? n : 1];
}

EdgeContainer::~EdgeContainer()
{
    delete[] target;
    delete[] source;
    delete[] weight;
}

Graph::Graph() {}

Graph::~Graph() { delete[] nodes; }

void Graph::allocateNodes(int size)
{
    if (nodes == nullptr) {
        nodes = new int[size > 0 ? size : 1];
        nodeCapacity = size;
    }
}

void Graph::allocateEdges(int size)
{
    if (edges.count == 0) {
        edges.allocate(size);
        edgeSize = 0;
    }
}

void Graph::addNode(int node)
{
    if (nodeSize >= nodeCapacity) throw std::out_of_range("Graph::addNode beyond allocateNodes capacity");
    nodes[nodeSize] = node;
    nodeMap[node] = nodeSize;
    ++nodeSize;
}

void Graph::addEdge(Edge edge)
{
    if (edgeSize >= edges.count) throw std::out_of_range("Graph::addEdge beyond allocateEdges capacity");
    edges.source[edgeSize] = edge.source;
    edges.target[edgeSize] = edge.target;
    edges.weight[edgeSize] = edge.weight;
    ++edgeSize;
}

void Graph::addEdgeFromVector(std::vector<Edge> list)
{
    for (const Edge &e : list) addEdge(e);
}

int Graph::setNodes(int *list, int size)
{
    delete[] nodes;
    nodes = new int[size > 0 ? size : 1];
    nodeCapacity = size;
    nodeSize = size;
    nodeMap.clear();
    for (int i = 0; i < size; ++i) {
        nodes[i] = list[i];
        nodeMap[list[i]] = i;
    }
    return 0;
}

int Graph::setEdges(Edge *list, int size)
{
    edges.allocate(size);
    edgeSize = size;
    for (int i = 0; i < size; ++i) {
        edges.source[i] = list[i].source;
        edges.target[i] = list[i].target;
        edges.weight[i] = list[i].weight;
    }
    return 0;
}

Edge Graph::getEdgeAt(int idx)
{
    Edge e;
    e.source = edges.source[idx];
    e.target = edges.target[idx];
    e.weight = edges.weight[idx];
    return e;
}

int Graph::getNodeAt(int idx) { return nodes[idx]; }
int *Graph::getNodes() { return nodes; }
EdgeContainer &Graph::getEdgeContainer() { return edges; }
int Graph::getNodeSize() { return nodeSize; }
int Graph::getEdgeSize() { return edgeSize; }

int Graph::getIdxOfNode(int node)
{
    auto it = nodeMap.find(node);
    return it == nodeMap.end() ? -1 : it->second;
}

void Graph::ensureNodeAttrs()
{
    if ((int)cluster_.size() != nodeSize) {
        cluster_.assign((size_t)nodeSize, 0);
        indexInCluster_.assign((size_t)nodeSize, 0);
        x_.assign((size_t)nodeSize, 0.f);
        y_.assign((size_t)nodeSize, 0.f);
    }
}

void Graph::setCluster(int idx, int cluster)
{
    ensureNodeAttrs();
    cluster_.at((size_t)idx) = cluster;
}
int Graph::getCluster(int idx)
{
    ensureNodeAttrs();
    return cluster_.at((size_t)idx);
}
void Graph::setIndexInCluster(int idx, int index)
{
    ensureNodeAttrs();
    indexInCluster_.at((size_t)idx) = index;
}
int Graph::getIndexInCluster(int idx)
{
    ensureNodeAttrs();
    return indexInCluster_.at((size_t)idx);
}
float Graph::getX(int idx)
{
    ensureNodeAttrs();
    return x_.at((size_t)idx);
}
float Graph::getY(int idx)
{
    ensureNodeAttrs();
    return y_.at((size_t)idx);
}
float *Graph::getXPtr(int idx)
{
    ensureNodeAttrs();
    return &x_.at((size_t)idx);
}
float *Graph::getYPtr(int idx)
{
    ensureNodeAttrs();
    return &y_.at((size_t)idx);
}

// The reference reads a count and then tokens until end of file; every token
// after the count is taken here (the reference drops the last node id when the
// file lacks trailing whitespace, LouvainCPP.cpp:676-692).
std::vector<int> loadNodesFile(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    long long count = 0;
    if (!(in >> count) || count < 0) throw std::runtime_error("bad node count in " + path);
    std::vector<int> ids;
    ids.reserve((size_t)count);
    std::string tok;
    while (in >> tok) ids.push_back(atoi(tok.c_str()));
    return ids;
}

std::vector<Edge> loadEdgesFile(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    long long count = 0;
    if (!(in >> count) || count < 0) throw std::runtime_error("bad edge count in " + path);
    std::vector<Edge> edges;
    edges.reserve((size_t)count);
    std::string a, b, w;
    while (in >> a) {
        if (!(in >> b >> w)) throw std::runtime_error("truncated edge triple in " + path);
        Edge e;
        e.source = atoi(a.c_str());
        e.target = atoi(b.c_str());
        e.weight = (float)atof(w.c_str());
        edges.push_back(e);
    }
    return edges;
}
