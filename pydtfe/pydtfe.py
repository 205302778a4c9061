"""The reference binding's functions (python/pydtfe.cpp:207-261):
compute_surface_density, compute_3d_density, interpolate_field2grid and
compute_particle_density, computed by libsdtfe_amd.so on the GPU."""
from sdtfe_amd.pydtfe import (compute_3d_density, compute_particle_density,  # noqa: F401
                              compute_surface_density, interpolate_field2grid)

__all__ = ["compute_surface_density", "compute_3d_density", "interpolate_field2grid",
           "compute_particle_density"]
