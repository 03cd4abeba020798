"""cbm -- host-side mirror of the CellBasedModels.jl API over libcbm_hip.so.

The reference's per-step hot path (``ABM`` / ``Community`` / ``evolve!``) runs
on an MI355X through the C ABI of ``include/cbm_hip.h``.  Names follow the
reference; Julia's mutating ``f!`` becomes ``f_`` here:

    abm = ABM(3, agentODE=CBMModels.repulsionAdhesion(k=1, r=1, eps=0.1),
              neighborsAlg=CBMNeighbors.VerletDisplacement(skin=2.7, nMaxNeighbors=32,
                                                          rebuildDisplacement=0.15),
              agentAlg=CBMIntegrators.Heun(), platform=HIP())
    com = Community(abm, N=1000, dt=0.005, simBox=[[0, 20], [0, 20], [0, 20]])
    com.x, com.y, com.z = ...
    evolve_(com, steps=100)

Reference: src/CellBasedModels.jl:15-97 (exports), src/platforms.jl:1-88
(platform seam), src/AgentStructure/agentStructure.jl:131-304 (ABM),
src/CommunityStructure/communityStructure.jl:71-187 (Community),
src/CommunityStructure/step.jl:96-155 (step!/evolve!).
"""
from ._abi import CBMError, lib  # noqa: F401
from .api import (  # noqa: F401
    HIP, CBMNeighbors, CBMIntegrators, CBMModels, CBMRules, ABM, Community,
    loadToPlatform_, bringFromPlatform_, computeNeighbors_, step_, evolve_,
    getParameter, version,
)
