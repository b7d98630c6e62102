"""hermeshd_amd -- MI355X-native discontinuous-Galerkin RHS and SSP-RK stage
update of HERMESHD, behind the reference's Python setup/step front end.

Importing the package does not touch the GPU; the HIP library is loaded on
first use (hermeshd_amd._lib).
"""
from .config import HDConfig, poiseuille_2d, sod_1d, vortex_3d  # noqa: F401

__all__ = ["HDConfig", "sod_1d", "vortex_3d", "poiseuille_2d"]
