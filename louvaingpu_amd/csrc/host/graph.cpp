// graph.cpp - host Graph/EdgeContainer storage (LouvainCUDA/Graph.cpp:28-160,
// 218-347, LouvainCUDA/Edge.h:26-63).  The CLI text loaders are in graphio.cpp.
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
