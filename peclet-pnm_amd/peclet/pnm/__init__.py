"""peclet.pnm — pore-network extraction from SDF pore geometry, on AMD MI355X (HIP, gfx950).

Drop-in for the reference package's module (packaging/pnm_init.py:7 re-exports ``_pnm``):
``extract_pores``, ``segment_volume``, ``extract_topology_gpu``, ``extract_pore_network``,
``extract_network_flow``, ``Pore``, ``SDFReader`` and ``execution_space``.
"""

from ._runtime import pin_hip_runtime as _pin

_pin()
from ._pnm import *  # noqa: F401,F403,E402
