"""haccsr -- MI355X-native short-range particle-particle force of HACC.

Python front end of libhaccsr.so.  The reference's public surface for this
path is one constructor, RCBForceTree<TDPTS>(...) (RCBForceTree.h:96-124),
which builds the tree, the interaction lists and applies the short-range
kick; ``forcetree.RCBMonopoleForceTree`` mirrors it over the C ABI, and
``_lib.Context`` exposes the phases separately.  ``particles.Particles``
mirrors the sub-cycle and PM particle-grid methods of the reference's
Particles class, and ``restart`` reads and writes its restart files.
"""
from ._lib import (Context, HsrError, LAW_NEWTON, LAW_SR_FIT, LAW_SR_INTERP,
                   LAW_SR_POLY5, LAW_SR_POLY6, LAW_STEP16, NODE_DTYPE, load)
from .forcetree import (FGrid, FGridEvalFit, FGridEvalInterp, FGridEvalPoly,
                        ForceLaw, ForceLawNewton, ForceLawSR, RCBMonopoleForceTree,
                        RCBQuadrupoleForceTree)

__all__ = [
    "Context", "HsrError", "LAW_NEWTON", "LAW_SR_FIT", "LAW_SR_INTERP",
    "LAW_SR_POLY5", "LAW_SR_POLY6", "LAW_STEP16", "NODE_DTYPE", "load", "FGrid",
    "FGridEvalFit", "FGridEvalInterp", "FGridEvalPoly", "ForceLaw", "ForceLawNewton", "ForceLawSR",
    "RCBMonopoleForceTree", "RCBQuadrupoleForceTree",
]
