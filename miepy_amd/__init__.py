"""miepy_amd: the MiePy cluster solve (generalized multiparticle Mie theory),
built for one AMD Instinct MI355X (gfx950).

The host side mirrors MiePy's `sphere_cluster` API and its backend dispatch
(`miepy_amd.backend` provides sphere_aggregate_tmatrix, solve_linear_system,
particle_cross_sections, force_all, torque_all with the reference's
signatures); the compute runs in hand-written HIP kernels in
lib/libmiepy_amd.so behind the C ABI of include/miepy_amd.h.
"""

from . import backends, materials as _materials_mod
from . import backend as _backend
from .backend import solver, DeviceTMatrix, cluster_coefficients, vsh_mode, vsh_translation
from .flux import cross_sections
from .materials import (constant_material, data_material, dielectric, function_material, material_struct,
                        materials)
from .sources import plane_wave, combined_source
from .sphere_cluster import sphere_cluster
from .cluster import cluster, sphere, sphere_cluster_particle, tmatrix_particle
from . import tmatrix

backends.register_backend("mi355x", _backend)


def lmax_to_rmax(lmax):
    return lmax * (lmax + 2)


def rmax_to_lmax(rmax):
    return int(-1 + (1 + rmax) ** 0.5)


def mode_indices(lmax):
    """(r, n, m) in MiePy's mode order (src/miepy/vsh/mode_indices.py:21-80)."""
    r = 0
    for n in range(1, lmax + 1):
        for m in range(-n, n + 1):
            yield r, n, m
            r += 1


__all__ = ["sphere_cluster", "cluster", "sphere", "tmatrix_particle", "plane_wave", "combined_source", "constant_material", "data_material", "dielectric",
           "function_material", "material_struct", "materials", "cross_sections", "solver", "DeviceTMatrix",
           "backends", "lmax_to_rmax", "rmax_to_lmax", "mode_indices", "cluster_coefficients", "vsh_mode",
           "vsh_translation"]
