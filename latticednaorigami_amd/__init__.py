"""latticednaorigami_amd - MI355X-native sampler for the LatticeDNAOrigami
lattice Monte Carlo model.

The product is libldo_hip.so (HIP kernels for gfx950 behind the C ABI of
include/ldo.h).  This package holds the host-side mirror of the reference's
simulation interface: system/moveset readers, the synthetic workloads and a
Python driver over the C ABI.
"""
from . import abi, systems  # noqa: F401
