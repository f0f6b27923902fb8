"""Scene and state I/O (opencloth_amd/io.py, oc_grid_triangles): the render
triangles of the reference (IE:473-488) checked by their geometry, the C and
Python versions against each other, and exact OBJ / NPZ round trips."""
import ctypes

import numpy as np

from opencloth_amd import io, scenes


def test_triangles_tile_the_grid():
    """Each quad is split into two triangles over its four corners along one
    diagonal, the diagonal alternates with the parity of row + column, all
    triangles of a flat grid face the same way and their areas sum to the
    grid's area."""
    nx, ny = 7, 5
    T = io.grid_triangles(nx, ny)
    assert T.shape == (2 * (nx - 1) * (ny - 1), 3)
    X = scenes.grid_positions(nx, ny, 4.0, 5.0).astype(np.float64)
    for q in range((nx - 1) * (ny - 1)):
        i, j = divmod(q, nx - 1)
        corners = {i * nx + j, i * nx + j + 1, (i + 1) * nx + j, (i + 1) * nx + j + 1}
        t1, t2 = set(T[2 * q]), set(T[2 * q + 1])
        assert t1 | t2 == corners and len(t1 & t2) == 2
        diag = t1 & t2
        main = {i * nx + j, (i + 1) * nx + j + 1}
        assert (diag == main) == ((i + j) % 2 == 0)
    n = np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]])
    assert (n[:, 1] > 0).all() or (n[:, 1] < 0).all()
    assert np.isclose(0.5 * np.abs(n[:, 1]).sum(), 4.0 * 4.0)


def test_c_abi_triangles_equal_python():
    from opencloth_amd._lib import lib, ptr

    for nx, ny in [(2, 2), (21, 21), (6, 9)]:
        out = np.zeros(6 * (nx - 1) * (ny - 1), np.int32)
        assert lib().oc_grid_triangles(nx, ny, ptr(out)) == 0
        assert np.array_equal(out.reshape(-1, 3), io.grid_triangles(nx, ny))
    assert lib().oc_grid_triangles(1, 5, ptr(np.zeros(1, np.int32))) == -1


def test_indices_past_16_bits():
    """512x512 has 262,144 vertices: indices stay exact (the reference's
    GLushort would wrap)."""
    T = io.grid_triangles(512, 512)
    assert T.max() == 512 * 512 - 1 and T.dtype == np.int32


def test_obj_round_trip(tmp_path):
    sc = scenes.large_scene(64)
    T = io.grid_triangles(sc.nx, sc.ny)
    p = str(tmp_path / "cloth.obj")
    io.write_obj(p, sc.X, T, comment="64x64 cloth\nperturbed")
    X, T2 = io.read_obj(p)
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(T2, T)


def test_npz_round_trip(tmp_path):
    sc = scenes.demo_scene()
    V = np.random.default_rng(0).standard_normal(sc.X.shape).astype(np.float32)
    p = str(tmp_path / "scene.npz")
    io.save_scene(p, sc, V=V)
    sc2 = io.load_scene(p)
    assert (sc2.nx, sc2.ny, sc2.steps, sc2.name) == (sc.nx, sc.ny, sc.steps, sc.name)
    assert sc2.params == sc.params
    assert np.array_equal(sc2.X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(sc2.V.view(np.uint32), V.view(np.uint32))
    assert sc2.springs.tobytes() == sc.springs.tobytes() and np.array_equal(sc2.pinned, sc.pinned)
    with np.load(p, allow_pickle=False) as z:
        assert np.array_equal(z["triangles"], io.grid_triangles(sc.nx, sc.ny))
