"""MI355X (gfx950) backend for Zog's `:soa` Brandes betweenness and PageRank.

``zog_amd.resource_graph`` mirrors ``Zog.ResourceGraph`` for the two
accelerated paths and calls the HIP library (``libzog_hip.so``) through the
C ABI in ``include/zog_hip.h``.
"""
from . import resource_graph  # noqa: F401
from .resource_graph import (  # noqa: F401
    ResourceGraph,
    betweenness_centrality,
    betweenness_unweighted,
    destroy,
    new,
    pagerank,
    read_edgelist,
)
from .soa import SoA  # noqa: F401
