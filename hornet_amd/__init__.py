"""hornet_amd -- MI355X (gfx950) implementation of Hornet's dynamic-graph hot path.

The compute lives in ``libhornet_hip.so`` (hand-written HIP, wave64) behind the
C ABI in ``include/hornet_hip.h``; this package is the host-side mirror of the
reference API (``Hornet``, ``BatchUpdate``, ``BfsTopDown``,
``TriangleCounting2``) over that ABI.
"""
from ._lib import HornetError, LIB_PATH, load  # noqa: F401


def __getattr__(name):
    # torch is imported lazily so that the ABI can be inspected without it
    if name in ("Hornet", "HornetInit", "BatchUpdate", "BatchUpdatePtr", "BfsTopDown", "BfsMulti",
                "TriangleCounting2", "StaticAlgorithm", "rmat_pairs_device", "chung_lu_pairs_device",
                "uniform_pairs_device", "INF"):
        from . import graph
        return getattr(graph, name)
    raise AttributeError(name)
