"""Loads libgf.so (built in-tree by csrc/Makefile) and declares its C ABI
(include/gf.h).  There is no fallback: if the library is missing or has no
GPU to run on, every call raises."""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgf.so")

_dbl = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_ci, _cd, _cf, _cll = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_longlong

# name -> (restype, argtypes)
SIGNATURES = {
    # reference engine ABI (cpp/nbody.cpp:415-505; engine.py:26-35)
    "nbody_num_threads": (_ci, []),
    "nbody_accel_active": (_ci, [_ci, _dbl, _dbl, _dbl, _cd, _cd, _vp, _ci]),
    "nbody_accel_direct": (_ci, [_ci, _dbl, _dbl, _dbl, _cd]),
    "nbody_render": (_ci, [_ci, _dbl, _u8, _ci, _dbl, _dbl, _cd, _ci, _ci, _f32]),
    "nbody_tonemap": (_ci, [_ci, _ci, _ci, _f32, _f32, _ci, _ci, _f32, _cf, _cf, _cf, _u8]),
    # library
    "gf_last_error": (ctypes.c_char_p, []),
    "gf_version": (_ci, []),
    "gf_device_info": (_ci, [ctypes.c_char_p, _ci]),
    "gf_set_precision": (_ci, [_ci]),
    "gf_set_test_limits": (_ci, [_ci, _ci, _ci]),
    "gf_set_quad_tolerance": (_ci, [_cd]),
    "gf_get_quad_tolerance": (_cd, []),
    "gf_set_quad_relative_tolerance": (_ci, [_cd]),
    "gf_get_quad_relative_tolerance": (_cd, []),
    "gf_set_sph_precision": (_ci, [_ci]),
    "gf_get_sph_precision": (_ci, []),
    "gf_set_knn_blocks": (_ci, [_ci]),
    "gf_get_knn_blocks": (_ci, []),
    "gf_set_knn_merge": (_ci, [_ci]),
    "gf_get_knn_merge": (_ci, []),
    "gf_set_walk_blocks": (_ci, [_ci]),
    "gf_get_walk_blocks": (_ci, []),
    "gf_get_precision": (_ci, []),
    "gf_last_interactions": (_cll, []),
    "gf_accel_direct_subset": (_ci, [_ci, _dbl, _dbl, _ci, _vp, _dbl, _cd]),
    "gf_tree_info": (_ci, [_vp, _vp, _vp]),
    "gf_lookback_timeouts": (_ci, []),
    "gf_tree_export": (_ci, [_ci] + [_vp] * 10),
    "gf_gravity_dev": (_ci, [_ci, _vp, _vp, _vp, _cd, _cd, _vp, _ci, _ci, _vp]),
    "gf_tree_build_dev": (_ci, [_ci, _vp, _vp, _cd, _vp]),
    "gf_walk_dev": (_ci, [_ci, _vp, _vp, _cd, _ci, _vp]),
    "gf_walk_prepare_dev": (_ci, [_ci, _vp]),
    "gf_direct_dev": (_ci, [_ci, _vp, _vp, _ci, _vp, _vp, _cd, _vp]),
    "gf_potential_dev": (_ci, [_ci, _vp, _vp, _vp, _cd, _vp]),
    "gf_combine_dev": (_ci, [_ci, _ci, _cd, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_step_drift_dev": (_ci, [_ci, _cd, _cll, _ci, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_step_kick_dev": (_ci, [_ci, _ci, _cll, _ci, _cd, _ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _cd, _vp]),
    "gf_interactions_dev": (_ci, [_vp]),
    "gf_node_interactions_dev": (_ci, [_vp]),
    "gf_sync": (_ci, [_vp]),
    # SPH
    "gf_sph_kernel": (_ci, [_ci, _dbl, _dbl, _vp, _vp]),
    "gf_sph_forces": (_ci, [_ci, _ci, _dbl, _dbl, _dbl, _i64, _dbl, _dbl, _cd, _cd, _cd,
                            _vp, _vp, _vp, _vp, _vp]),
    "gf_knn": (_ci, [_ci, _dbl, _ci, _vp, _i64, _dbl, _vp]),
    "gf_knn_from_candidates_host": (_ci, [_ci, _ci, _dbl, _i64, _ci, _i64, _dbl]),
    "gf_nbcache_search": (_ci, [_ci, _ci, _i64, _dbl, _ci, _ci, _i64, _dbl]),
    "gf_nbcache_reselect": (_ci, [_ci, _ci, _i64, _dbl, _ci, _i64, _dbl]),
    "gf_nbcache_info": (_ci, [_vp, _vp, _vp]),
    "gf_nbcache_export": (_ci, [_vp, _vp]),
    "gf_hydro_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _cd, _ci, _vp, _cd, _cd, _cd,
                           _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_hydro_setup_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _vp]),
    "gf_hydro_setup_from_tree_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _vp]),
    "gf_hydro_knn_dev": (_ci, [_ci, _vp, _vp]),
    "gf_hydro_knn_begin_dev": (_ci, [_ci, _vp, _ci, _vp]),
    "gf_hydro_knn_more_dev": (_ci, [_ci, _vp]),
    "gf_hydro_knn_end_dev": (_ci, [_vp]),
    "gf_hydro_forces_dev": (_ci, [_cd, _cd, _cd, _cd, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_hydro_passes": (_ci, [_vp, _vp]),
    "gf_hydro_dk_dev": (_ci, [_vp, _vp]),
    "gf_hydro_stats": (_ci, [_i64, _ci]),
    # learned smoothing length
    "gf_mlp_params": (_ci, []),
    "gf_hydro_cell_density_dev": (_ci, [_ci, _vp, _vp, _vp]),
    "gf_mlp_features_dev": (_ci, [_ci, _ci, _vp, _vp, _vp, _vp, _vp, _cf, _vp, _vp, _vp, _vp]),
    "gf_mlp_forward_dev": (_ci, [_ci, _vp, _vp, _vp, _vp]),
    "gf_mlp_propose_dev": (_ci, [_ci, _vp, _vp, _vp, _cf, _vp, _vp]),
    "gf_mlp_grad_dev": (_ci, [_ci, _vp, _vp, _vp, _cf, _vp, _vp, _vp]),
    "gf_mlp_train_step_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _vp, _ci, _cf, _cf, _cf, _cf, _cf,
                                    _vp, _vp]),
    "gf_mlp_train_step_noalloc_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _vp, _ci, _cf, _cf, _cf, _cf,
                                            _cf, _vp, _vp]),
    "gf_mlp_grad_host": (_ci, [_ci, np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"), _f32,
                               _f32, _cf, _f32, _vp]),
    "gf_hydro_neighbours": (_ci, [_i64, _vp]),
    # gas physics (cooling, thermodynamics, star formation, feedback)
    "gf_gas_constants": (_ci, [_dbl]),
    "gf_cooling_function_dev": (_ci, [_ci, _vp, _vp, _vp]),
    "gf_cool_dev": (_ci, [_ci, _vp, _vp, _vp, _cd, _vp, _cd, _vp, _vp]),
    "gf_thermo_dev": (_ci, [_ci, _vp, _ci, _vp, _vp, _vp, _vp, _cd, _cd, _cd, _cd, _vp]),
    "gf_gas_errors_dev": (_ci, [_vp]),
    "gf_sf_candidates_dev": (_ci, [_ci, _vp, _ci, _vp, _vp, _cd, _cd, _cd, _cd, _vp, _vp, _vp]),
    "gf_sf_select_dev": (_ci, [_ci, _vp, _vp, _vp, _vp, _vp]),
    "gf_knn_query_dev": (_ci, [_ci, _vp, _ci, _vp, _ci, _vp, _vp, _vp]),
    "gf_feedback_heat_dev": (_ci, [_ci, _ci, _vp, _vp, _cd, _cd, _ci, _vp, _vp, _vp, _vp]),
}

_lib = None


def _preload_torch_hip_runtime():
    """One HIP runtime per process.  libgf.so needs libamdhip64.so.7; PyTorch
    ships its own copy (same SONAME) and loads it under another file name.
    Loading PyTorch's copy first, by path, makes libgf.so and a later
    `import torch` share it; without PyTorch, /opt/rocm's is used."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


class GFError(RuntimeError):
    """A libgf.so entry point failed (message from gf_last_error)."""


def load():
    """The library, with argtypes declared.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GFError(f"{LIB_PATH} is missing: build it first (python -c "
                          "'import __graft_entry__ as g; g.build()' or make -C "
                          "galaxy_formation_amd/csrc)")
        _preload_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Raises GFError with the library's message for a negative status."""
    if rc < 0:
        msg = load().gf_last_error()
        raise GFError(msg.decode() if msg else f"libgf call failed ({rc})")
    return rc


def ptr(a):
    """ctypes pointer of a numpy array (None stays None)."""
    return None if a is None else a.ctypes.data_as(_vp)
