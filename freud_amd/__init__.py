"""freud_amd: freud's pair-analysis path (neighbour queries, RDF, Steinhardt,
SolidLiquid, Hexatic, Cluster) on one AMD Instinct MI355X.

The Python API mirrors freud 3.5.0's ``freud.box``, ``freud.locality``,
``freud.density``, ``freud.environment`` (BondOrder), ``freud.order``, ``freud.pmft`` and ``freud.cluster``
for this path; the work
runs in hand-written HIP kernels (``freud_amd/csrc``) behind the C-ABI declared
in ``include/freud_amd.h``.
"""

from freud_amd import (batch, box, cluster, data, dcd, density, environment, gsd, learned, locality, order,
                       parallel, pmft)
from freud_amd.box import Box
from freud_amd.locality import AABBQuery, CellQuery, LinkCell, NeighborList

__version__ = "0.1.0"

__all__ = [
    "AABBQuery",
    "Box",
    "batch",
    "CellQuery",
    "LinkCell",
    "NeighborList",
    "box",
    "cluster",
    "data",
    "density",
    "environment",
    "learned",
    "locality",
    "order",
    "parallel",
    "pmft",
]
