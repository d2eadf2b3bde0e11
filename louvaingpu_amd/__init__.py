"""MI355X-native Louvain community detection (drop-in for LouvainGPU's GPU path).

The device engine is ``liblouvain_amd.so`` (hand-written HIP for gfx950, C ABI
in ``include/louvain_amd.h``); the host layer ``liblouvain_host.so`` keeps the
reference's Graph/Partition API and level loop.  Python reaches both through
ctypes; there is no host fallback for any phase.
"""
from .engine import Engine, generate_dendogram_phases, make_config  # noqa: F401
from .batch import GraphSet  # noqa: F401
from .serve import Server  # noqa: F401
from .louvain import Dendrogram, count_partitions, generate_dendogram, start  # noqa: F401
from . import generators  # noqa: F401

__all__ = ["Engine", "GraphSet", "Server", "generate_dendogram_phases", "make_config", "Dendrogram",
           "count_partitions", "generate_dendogram", "start", "generators"]
