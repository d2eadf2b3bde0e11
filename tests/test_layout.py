"""Cluster sizes and circle layout (calculateClusterSize + layoutProcess,
LouvainCUDA/LouvainCPP.cpp:555-651): the host library against a literal
restatement of the reference's loops (pure Python, libm cos/sin in double,
stored as float32), bit for bit; the device version against the host."""
import math

import numpy as np
import pytest

from conftest import gpu_case
from louvaingpu_amd import Engine
from louvaingpu_amd.louvain import cluster_layout


def reference_layout(ids, node, comm, nclusters):
    """The reference's loops as written (nested scan for the cluster, running
    counts, prefix sums, then the per-cluster layout of the 8 threads)."""
    n = len(ids)
    size = [0] * nclusters
    cluster = [0] * n
    index = [0] * n
    for i in range(n):
        for j in range(len(node)):
            if node[j] == ids[i]:
                c = comm[j]
                break
        cluster[i] = c
        index[i] = size[c]
        size[c] += 1
    for i in range(1, nclusters):
        size[i] += size[i - 1]
    x = [0.0] * n
    y = [0.0] * n
    for j in range(nclusters):
        metr = size[j - 1] if j > 0 else 0
        for i in range(n):
            if cluster[i] == j:
                s = metr + index[i]
                x[i] = np.float32(50 * math.cos(2 * math.pi * s / n))
                y[i] = np.float32(50 + 50 * math.sin(2 * math.pi * s / n))
    return (np.array(size, np.int32), np.array(cluster, np.int32), np.array(index, np.int32),
            np.array(x, np.float32), np.array(y, np.float32))


def _case(seed, n, k):
    rng = np.random.Generator(np.random.PCG64(seed))
    ids = rng.permutation(np.arange(n) * 3 + 7).astype(np.int32)
    comm_of = rng.integers(0, k, n).astype(np.int32)
    order = rng.permutation(n)  # the Partition lists nodes in its own order
    return ids, ids[order], comm_of[order], k


@pytest.mark.parametrize("seed,n,k", [(0, 1, 1), (1, 10, 3), (2, 57, 9), (3, 300, 40),
                                      (4, 200, 1), (5, 128, 128)])
def test_host_layout_matches_reference_loops(seed, n, k):
    ids, node, comm, k = _case(seed, n, k)
    got = cluster_layout(ids, node, comm, k)
    exp = reference_layout(ids.tolist(), node.tolist(), comm.tolist(), k)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_host_layout_rejects_missing_node():
    from louvaingpu_amd._lib import LouvainError, LV_ERR_INVALID_INPUT
    with pytest.raises(LouvainError) as ei:
        cluster_layout(np.array([1, 2], np.int32), np.array([1], np.int32),
                       np.array([0], np.int32), 1)
    assert ei.value.status == LV_ERR_INVALID_INPUT


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,k", [(1, 10, 3), (3, 300, 40), (6, 200000, 3000)])
def test_device_layout_matches_host(seed, n, k):
    gpu_case("device layout n=%d k=%d" % (n, k))
    ids = np.arange(n, dtype=np.int32)
    comm = np.random.Generator(np.random.PCG64(seed)).integers(0, k, n).astype(np.int32)
    cs, cl, ix, x, y = cluster_layout(ids, ids, comm, k)
    with Engine() as eng:
        ce, dix, dx, dy = eng.cluster_layout(comm, k)
    assert np.array_equal(ce, cs)
    assert np.array_equal(dix, ix)
    # device cos/sin (double) may differ from libm in the last bit of the double,
    # which can move the float32 result by one ulp
    assert np.all(np.abs(dx.view(np.int32) - x.view(np.int32)) <= 1)
    assert np.all(np.abs(dy.view(np.int32) - y.view(np.int32)) <= 1)
