"""autojax_amd: MI355X-native interferometer-inversion likelihood.

Modules
-------
api        drop-in functions with the names and parameters of autojax.numba
evidence   batched log-evidence over many samples sharing one W-tilde
datasets   synthetic workloads (grids, uv coverage, Delaunay sources, mappings)
build      hipcc recipe for lib/libautojax_amd.so (gfx950)
"""

__version__ = "0.1.0"
