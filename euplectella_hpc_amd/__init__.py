"""MI355X-native D3Q19 BGK lattice-Boltzmann hot path for the sponge-flow solver.

The compute path is hand-written HIP for gfx950 behind the C-ABI declared in ``include/eup.h``;
``euplectella_hpc_amd.lib`` binds it with ctypes.  Geometry generators, reference file formats
and the D3Q19 tables are host utilities.
"""
__all__ = ["d3q19", "fortio", "geometry"]
