"""MI355X-native engine for AxBC's sampled-Brandes hot path.

    from axbc_amd import Graph, Engine
    g = Graph.kronecker(16, 16, seed=(1, 2))
    with Engine(g) as e:
        e.pivot(v0, 1.0 / g.N)
        bc = e.bc()
"""
from .graph import Graph, kronecker_edge_list  # noqa: F401
from .engine import Engine, EngineError, device_count  # noqa: F401
