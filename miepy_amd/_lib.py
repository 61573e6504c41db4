"""ctypes binding of libmiepy_amd.so (the C ABI declared in include/miepy_amd.h).

The library is loaded from miepy_amd/lib/ only; there is no fallback path: a
missing library or a machine without a GPU raises immediately.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmiepy_amd.so")

MPA_OK = 0
MPA_ERR_ARG = 1
MPA_ERR_HIP = 2
MPA_NOT_CONVERGED = 3
MPA_BREAKDOWN = 4
MPA_NONFINITE = 5
MPA_ERR_OOM = 6

MPA_C128 = 0
MPA_C64 = 1
MPA_OPT_GRAPHS = 0
MPA_OPT_MATFREE_FP64_KERNEL = 1
MPA_OPT_MATFREE_C64_KERNEL = 2
MPA_OPT_OVERLAP = 3
MPA_SOLVE_REF_LOOP = 1

STATUS_NAMES = {
    MPA_OK: "ok",
    MPA_ERR_ARG: "bad argument",
    MPA_ERR_HIP: "HIP error",
    MPA_NOT_CONVERGED: "not converged",
    MPA_BREAKDOWN: "breakdown",
    MPA_NONFINITE: "non-finite",
    MPA_ERR_OOM: "out of memory",
}


class MpaError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}" + (f" ({detail})" if detail else ""))


_i = ctypes.c_int
_d = ctypes.c_double
_v = ctypes.c_void_p
_sz = ctypes.c_size_t
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)
_pv = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); kept in the order of include/miepy_amd.h
SIGNATURES = {
    "mpa_version": (ctypes.c_char_p, []),
    "mpa_last_error": (ctypes.c_char_p, []),
    "mpa_sphere_aggregate_tmatrix": (_i, [_i, _i, _v, _v, _d, _i, _i, _v, _v]),
    "mpa_sphere_aggregate_tmatrix_dual": (_i, [_i, _i, _v, _v, _d, _i, _v, _v, _v]),
    "mpa_matvec": (_i, [_i, _i, _i, _v, _i, _v, _v, _v]),
    "mpa_bicgstab_workspace_bytes": (_sz, [_i, _i, _i]),
    "mpa_bicgstab": (_i, [_i, _i, _i, _v, _i, _v, _v, _d, _i, _v, _pi, _pd, _pi, _i, _v]),
    "mpa_bicgstab_noalloc": (_i, [_i, _i, _i, _v, _i, _v, _v, _d, _i, _pi, _pd, _pi, _v]),
    "mpa_matfree_matvec_batched": (_i, [_i, _v, _pi, _i, _i, _v, _v, _v]),
    "mpa_bicgstab_ex_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "mpa_lu_workspace_bytes": (_sz, [_i]),
    "mpa_mie_coefficients": (_i, [_i, _i, _v, _v, _v, _v, _v, _v, _v, _v, _v, _v]),
    "mpa_plane_wave_structure": (_i, [_i, _i, _i, _v, _d, _v, _v, _v, _v, _v, _v]),
    "mpa_cluster_coefficients_workspace_bytes": (_sz, [_i, _i, _i]),
    "mpa_cluster_coefficients": (_i, [_i, _i, _i, _i, _v, _d, _d, _d, _v, _d, _v, _v, _v]),
    "mpa_cluster_source": (_i, [_i, _i, _i, _v, _d, _d, _d, _d, _v, _v]),
    "mpa_cluster_tmatrix": (_i, [_i, _i, _i, _v, _d, _d, _d, _v, _v, _d, _v, _v, _v]),
    "mpa_cluster_cross_sections": (_i, [_i, _i, _v, _v, _d, _v, _v]),
    "mpa_vsh_translation": (_i, [_i, _v, _i, _i, _v, _v, _v, _d, _i, _v, _v]),
    "mpa_lu_factor": (_i, [_i, _v, _i, _v, _pi, _v]),
    "mpa_lu_solve": (_i, [_i, _i, _v, _v, _v, _v, _v]),
    "mpa_bicgstab_ex": (_i, [_i, _i, _i, _v, _i, _v, _v, _d, _i, _i, _v, _i, _v, _pi, _pd, _pi, _i, _v]),
    "mpa_refine_workspace_bytes": (_sz, [_i, _i]),
    "mpa_solve_refined": (_i, [_i, _i, _v, _i, _v, _i, _v, _v, _d, _d, _i, _i, _v, _pi, _pi, _pd, _pi, _v]),
    "mpa_cluster_observables": (_i, [_i, _i, _i, _v, _v, _v, _v, _d, _d, _v, _v, _v, _v, _v, _v, _v]),
    "mpa_force_torque": (_i, [_i, _i, _v, _v, _d, _d, _v, _v, _v]),
    "mpa_particle_cross_sections": (_i, [_i, _i, _v, _v, _v, _d, _v, _v]),
    "mpa_stream_copy": (_i, [_v, _v, _sz, _v]),
    "mpa_autotune_matvec": (_i, [_i, _i, _i, _v, _v, _v, _i, _pi, _pd, _v]),
    "mpa_autotune_matvec_batched": (_i, [_i, _pi, _pi, _i, _pv, _pv, _pv, _i, _pi, _pd, _v]),
    "mpa_autotune_assembly": (_i, [_i, _i, _v, _v, _d, _i, _v, _i, _pi, _pd, _v]),
    "mpa_assembly_variant": (_i, [_i, _i, _pi]),
    "mpa_matvec_variant": (_i, [_i, _i, _i]),
    "mpa_matvec_variant_name": (ctypes.c_char_p, [_i, _i, _i]),
    "mpa_set_option": (_i, [_i, _i]),
    "mpa_get_option": (_i, [_i]),
    "mpa_profile_enable": (_i, [_i]),
    "mpa_profile_read": (_i, [_i, _pi, _pd]),
    "mpa_translation_table": (_i, [_i, _pi, _pd]),
    "mpa_assembly_tile": (_i, [_i]),
}

class MpaSystem(ctypes.Structure):
    """mpa_system (include/miepy_amd.h)"""
    _fields_ = [("n", ctypes.c_int), ("nrhs", ctypes.c_int), ("add_identity", ctypes.c_int), ("flags", ctypes.c_int),
                ("A", ctypes.c_void_p), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p), ("state", ctypes.c_void_p),
                ("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("block_size", ctypes.c_int),
                ("op", ctypes.c_void_p), ("minv", ctypes.c_void_p)]


class MpaClusterOperator(ctypes.Structure):
    """mpa_cluster_operator (include/miepy_amd.h)"""
    _fields_ = [("N", ctypes.c_int), ("lmax", ctypes.c_int), ("positions", ctypes.c_void_p),
                ("mie", ctypes.c_void_p), ("k", ctypes.c_double), ("workspace", ctypes.c_void_p),
                ("tmatrix", ctypes.c_void_p)]


class MpaRefineSystem(ctypes.Structure):
    """mpa_refine_system (include/miepy_amd.h)"""
    _fields_ = [("n", ctypes.c_int), ("nrhs", ctypes.c_int), ("add_identity", ctypes.c_int), ("pad", ctypes.c_int),
                ("A32", ctypes.c_void_p), ("A64", ctypes.c_void_p), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("state", ctypes.c_void_p), ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("maxiter", ctypes.c_int), ("max_refine", ctypes.c_int), ("op", ctypes.c_void_p)]


class MpaAsmCluster(ctypes.Structure):
    """mpa_asm_cluster (include/miepy_amd.h)"""
    _fields_ = [("N", ctypes.c_int), ("lmax", ctypes.c_int), ("dtype", ctypes.c_int), ("add_identity", ctypes.c_int),
                ("positions", ctypes.c_void_p), ("mie", ctypes.c_void_p), ("k", ctypes.c_double),
                ("out", ctypes.c_void_p), ("out32", ctypes.c_void_p)]


class MpaObsCluster(ctypes.Structure):
    """mpa_obs_cluster (include/miepy_amd.h)"""
    _fields_ = [("N", ctypes.c_int), ("lmax", ctypes.c_int), ("nsrc", ctypes.c_int), ("pad", ctypes.c_int),
                ("p_inc", ctypes.c_void_p), ("p_src", ctypes.c_void_p), ("mie_scat", ctypes.c_void_p),
                ("mie_int", ctypes.c_void_p), ("k", ctypes.c_double), ("eps_b", ctypes.c_double),
                ("p_scat", ctypes.c_void_p), ("p_int", ctypes.c_void_p), ("C_part", ctypes.c_void_p),
                ("C_total", ctypes.c_void_p), ("F", ctypes.c_void_p), ("T", ctypes.c_void_p)]


_ps = ctypes.POINTER(MpaSystem)
_pop = ctypes.POINTER(MpaClusterOperator)
_prs = ctypes.POINTER(MpaRefineSystem)
_pobs = ctypes.POINTER(MpaObsCluster)
_pasm = ctypes.POINTER(MpaAsmCluster)
SIGNATURES.update({
    "mpa_particle_aggregate_tmatrix": (_i, [_i, _i, _v, _v, _d, _i, _i, _v, _v]),
    "mpa_tmatrix_apply": (_i, [_i, _i, _i, _v, _v, _v, _v]),
    "mpa_matfree_workspace_bytes": (_sz, [_i, _i, _i]),
    "mpa_matfree_matvec": (_i, [_pop, _i, _i, _v, _v, _v]),
    "mpa_matfree_matvec_c64": (_i, [_pop, _i, _i, _v, _v, _v]),
    "mpa_autotune_matfree": (_i, [_pop, _i, _i, _v, _v, _i, _pi, _pd, _v]),
    "mpa_matfree_variant": (_i, [_i, _i, _pi]),
    "mpa_refine_systems_start": (_i, [_i, _prs, _v, _v]),
    "mpa_refine_systems_advance": (_i, [_i, _prs, _v, _i, _pi, _pi, _pi, _pd, _pi, _v]),
    "mpa_bicgstab_matfree": (_i, [_pop, _i, _i, _v, _v, _d, _i, _v, _pi, _pd, _pi, _i, _v]),
    "mpa_system_state_bytes": (_sz, [_i, _i, _i]),
    "mpa_system_state_bytes_ex": (_sz, [_i, _i, _i, _i]),
    "mpa_batch_scratch_bytes": (_sz, [_i, _ps, _i]),
    "mpa_bicgstab_batched": (_i, [_i, _ps, _i, _v, _pi, _pd, _pi, _i, _v]),
    "mpa_systems_start": (_i, [_i, _ps, _i, _v, _v]),
    "mpa_systems_advance": (_i, [_i, _ps, _i, _v, _i, _pi, _pi, _pd, _pi, _v]),
    "mpa_refine_state_bytes": (_sz, [_i, _i]),
    "mpa_refine_scratch_bytes": (_sz, [_i, _prs]),
    "mpa_solve_refined_batched": (_i, [_i, _prs, _v, _pi, _pi, _pd, _pi, _v]),
    "mpa_cluster_observables_batched": (_i, [_i, _pobs, _v]),
    "mpa_sphere_aggregate_tmatrix_batched": (_i, [_i, _pasm, _v]),
})

_lib = None


def load():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m miepy_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error():
    return load().mpa_last_error().decode(errors="replace")


def check(status, where):
    if status != MPA_OK:
        raise MpaError(status, where, last_error())
    return status
