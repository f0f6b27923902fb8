"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares, with a Python signature for each."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(oc_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "oc_cloth_create" in names and "oc_cloth_step" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from opencloth_amd._lib import LIB_PATH, SIGNATURES, lib

    L = lib()  # loads; no GPU call is made
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    unbound = [n for n in declared_functions() if n not in SIGNATURES]
    assert not unbound, unbound
    assert os.path.exists(LIB_PATH)


def test_version_string():
    from opencloth_amd._lib import lib

    assert b"gfx950" in lib().oc_version()


def test_create_rejects_bad_input_without_device():
    """Argument errors come back as status codes (no exception, no crash)."""
    from opencloth_amd._lib import OcParams, lib

    h = ctypes.c_void_p()
    p = OcParams()
    p.mass = 1.0
    p.dt = 0.0  # invalid
    rc = lib().oc_cloth_create(3, 3, None, None, None, 0, None, ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert lib().oc_last_error() != b""


def test_pbd_create_rejects_bad_constraints_without_device():
    """oc_pbd_create checks particle indices and parameters before touching
    the device."""
    import numpy as np

    from opencloth_amd import scenes
    from opencloth_amd._lib import OC_ERR_INVALID_ARGUMENT, OcPbdParams, lib, ptr

    X = np.zeros((4, 3), np.float32)
    d = np.zeros(1, scenes.PBD_DISTANCE_DTYPE)
    d["p1"], d["p2"], d["k"] = 0, 4, 0.5  # index out of range
    b = np.zeros(0, scenes.PBD_BENDING_DTYPE)
    p = OcPbdParams()
    p.mass, p.dt, p.solver_iterations = 1.0, 0.01, 2
    h = ctypes.c_void_p()
    rc = lib().oc_pbd_create(4, ptr(X), None, ptr(d), 1, ptr(b), 0, ctypes.byref(p), ctypes.byref(h))
    assert rc == OC_ERR_INVALID_ARGUMENT and b"distance constraint 0" in lib().oc_last_error()
    d["p2"] = 1
    p.solver_iterations = 0
    rc = lib().oc_pbd_create(4, ptr(X), None, ptr(d), 1, ptr(b), 0, ctypes.byref(p), ctypes.byref(h))
    assert rc == OC_ERR_INVALID_ARGUMENT and not h.value


def test_tile_shape_is_reported_without_device():
    """oc_cloth_tile_shape (host only) gives the shape the oracle must use for
    the kernels' reduction order: whole waves per tile."""
    from opencloth_amd._lib import tile_shape

    tx, ty = tile_shape()
    assert tx > 0 and ty > 0 and (tx * ty) % 64 == 0 and tx * ty <= 1024


def test_vgrid_create_rejects_bad_input_without_device():
    """oc_vgrid_create checks the grid, the positions and the parameters
    before touching the device; NULL handles come back as errors."""
    import numpy as np

    from opencloth_amd import scenes
    from opencloth_amd._lib import OC_ERR_INVALID_ARGUMENT, OcVgridParams, lib, ptr

    X = np.zeros((4, 4), np.float32)
    good = scenes.vgrid_params(2, 2)
    h = ctypes.c_void_p()
    for w, hgt, x, change in ((0, 4, X, {}), (2, 2, None, {}), (2, 2, X, {"dt": 0.0}),
                              (2, 2, X, {"mass": -1.0}), (1 << 15, 1 << 14, X, {})):
        p = OcVgridParams.from_dict(dict(good, **change))
        rc = lib().oc_vgrid_create(w, hgt, ptr(x), None, None, ctypes.byref(p), ctypes.byref(h))
        assert rc == OC_ERR_INVALID_ARGUMENT and not h.value, (w, hgt, change)
        assert lib().oc_last_error() != b""
    assert lib().oc_vgrid_step(None, 1, None) == OC_ERR_INVALID_ARGUMENT
    assert lib().oc_vgrid_destroy(None) == 0
