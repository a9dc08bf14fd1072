"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function include/gf.h declares, with the reference engine's five symbols
(cpp/nbody.cpp:415-505) among them."""

import ctypes
import os
import re

from galaxy_formation_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "gf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|long long|const char\*|void)\s+\**(\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 5
    for name in names:
        assert hasattr(lib, name), name


def test_reference_engine_symbols_present():
    names = declared()
    for s in ("nbody_num_threads", "nbody_accel_active", "nbody_accel_direct", "nbody_render",
              "nbody_tonemap"):
        assert s in names


def test_python_signatures_cover_the_header():
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_loader_declares_argtypes():
    L = _lib.load()
    assert L.nbody_accel_active.argtypes is not None
    assert L.gf_version() >= 1


def test_knn_blocks_setting_is_validated():
    """The neighbour pass's block count: default 1280, 0 restores it,
    negative values are refused (no GPU needed)."""
    L = _lib.load()
    assert L.gf_get_knn_blocks() == 1280
    assert L.gf_set_knn_blocks(-1) < 0
    assert L.gf_set_knn_blocks(512) == 0 and L.gf_get_knn_blocks() == 512
    assert L.gf_set_knn_blocks(0) == 0 and L.gf_get_knn_blocks() == 1280


def test_quad_relative_tolerance_setting_is_validated():
    """The relative quadrupole rule: off (0) by default, values outside
    [0, 1] and NaN are refused and leave the setting as it was (no GPU
    needed)."""
    L = _lib.load()
    assert L.gf_get_quad_relative_tolerance() == 0.0
    assert L.gf_set_quad_relative_tolerance(-1e-4) < 0
    assert L.gf_set_quad_relative_tolerance(float("nan")) < 0
    assert L.gf_set_quad_relative_tolerance(2.0) < 0
    assert L.gf_get_quad_relative_tolerance() == 0.0
    assert L.gf_set_quad_relative_tolerance(1e-4) == 0
    assert abs(L.gf_get_quad_relative_tolerance() - 1e-4) < 1e-4 * 1e-6   # held as FP32
    assert L.gf_set_quad_relative_tolerance(0.0) == 0 and L.gf_get_quad_relative_tolerance() == 0.0


def test_engine_rejects_misshapen_arrays():
    """Engine.gravity refuses pos/acc that are not (n, 3) and an active mask
    of another length before any library call (the library would read and
    write 3 n doubles)."""
    import numpy as np
    import pytest
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp32")
    pos, mass = np.zeros((10, 3)), np.ones(10)
    with pytest.raises(ValueError):
        e.gravity(pos, mass, np.zeros((10, 1)), 0.5, 0.01)
    with pytest.raises(ValueError):
        e.gravity(np.zeros((9, 3)), mass, np.zeros((10, 3)), 0.5, 0.01)
    with pytest.raises(ValueError):
        e.gravity(pos, mass, np.zeros((10, 3)), 0.5, 0.01, active=np.ones(9, np.uint8))
    with pytest.raises(ValueError):
        e.gravity_direct(np.zeros((10, 2)), mass, 0.01)


def test_sph_precision_setting_is_validated():
    """SPH pair precision: 0 (FP64) by default, 1 = FP32 pair terms; other
    values are refused and leave the setting (no GPU needed)."""
    L = _lib.load()
    assert L.gf_get_sph_precision() == 0
    assert L.gf_set_sph_precision(2) < 0 and L.gf_set_sph_precision(-1) < 0
    assert L.gf_get_sph_precision() == 0
    assert L.gf_set_sph_precision(1) == 0 and L.gf_get_sph_precision() == 1
    assert L.gf_set_sph_precision(0) == 0 and L.gf_get_sph_precision() == 0


def test_walk_blocks_setting_is_validated():
    """The walk's persistent grid: default 4096, 0 restores it, values
    outside [0, 65536] are refused (no GPU needed)."""
    L = _lib.load()
    assert L.gf_get_walk_blocks() == 4096
    assert L.gf_set_walk_blocks(-1) < 0 and L.gf_set_walk_blocks(65537) < 0
    assert L.gf_set_walk_blocks(3072) == 0 and L.gf_get_walk_blocks() == 3072
    assert L.gf_set_walk_blocks(0) == 0 and L.gf_get_walk_blocks() == 4096
