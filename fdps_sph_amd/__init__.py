"""fdps_sph_amd -- MI355X (gfx950) implementation of the FDPS_SPH per-step
interaction path: SPH density / derivative / hydro-force and self-gravity
functors, EoS table evaluation, and a device-resident step driver, behind the
C-ABI of include/fdps_sph_gpu.h."""
from ._lib import LIB_PATH, SphGpuError, lib  # noqa: F401
from .sph import Context, Simulation, default_params  # noqa: F401

__all__ = ["Context", "Simulation", "default_params", "SphGpuError", "lib", "LIB_PATH"]
