"""Scene and state I/O: the grid's triangle list, OBJ export/import of the
cloth surface and NPZ snapshots of a whole scene.

The reference builds its render triangles in InitGL
(OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp:473-488): two
triangles per quad, the diagonal alternating with the parity of (row +
column).  It stores them as GLushort, which overflows past 65,535 vertices
(IE:276); here indices are 32-bit.  OBJ files hold positions with nine
significant digits, enough to read every float32 back exactly.  NPZ
snapshots store plain and structured numpy arrays only (no pickled
objects), so they load with numpy's default allow_pickle=False.
"""
from __future__ import annotations

import dataclasses
import json

import numpy as np

from .scenes import SPRING_DTYPE, Scene, StepParams


def grid_triangles(nx: int, ny: int) -> np.ndarray:
    """(2*(nx-1)*(ny-1), 3) int32 vertex indices in the reference's order:
    quad (row i, column j) with corners i0 = i*nx + j, i1 = i0 + 1,
    i2 = i0 + nx, i3 = i2 + 1 gives (i0, i2, i1), (i1, i2, i3) when i + j is
    odd and (i0, i2, i3), (i0, i3, i1) when it is even (IE:473-488)."""
    if nx < 2 or ny < 2:
        return np.zeros((0, 3), np.int32)
    i, j = np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    i0 = i * nx + j
    i1, i2 = i0 + 1, i0 + nx
    i3 = i2 + 1
    odd = ((i + j) % 2 == 1)[..., None]
    t_odd = np.stack([np.stack([i0, i2, i1], -1), np.stack([i1, i2, i3], -1)], -2)
    t_even = np.stack([np.stack([i0, i2, i3], -1), np.stack([i0, i3, i1], -1)], -2)
    tris = np.where(odd[..., None], t_odd, t_even)
    return tris.reshape(-1, 3).astype(np.int32)


def write_obj(path: str, X: np.ndarray, triangles: np.ndarray, comment: str | None = None) -> None:
    """Wavefront OBJ: one `v x y z` per vertex, one `f a b c` (1-based) per
    triangle."""
    X = np.asarray(X, np.float32)
    T = np.asarray(triangles, np.int64)
    with open(path, "w") as f:
        if comment:
            for line in comment.splitlines():
                f.write(f"# {line}\n")
        f.write("".join("v %.9g %.9g %.9g\n" % (x, y, z) for x, y, z in X.tolist()))
        f.write("".join("f %d %d %d\n" % (a + 1, b + 1, c + 1) for a, b, c in T.tolist()))


def read_obj(path: str):
    """Positions (float32) and triangles (int32, 0-based) of an OBJ file;
    faces with more than three corners are fanned, texture and normal
    indices (`f a/b/c`) are ignored."""
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] == "v":
                verts.append([float(p) for p in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) for p in parts[1:]]
                idx = [k - 1 if k > 0 else len(verts) + k for k in idx]
                for k in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(verts, np.float32).reshape(-1, 3), np.asarray(faces, np.int32).reshape(-1, 3)


def save_scene(path: str, sc: Scene, X=None, V=None) -> None:
    """NPZ snapshot of a scene, optionally with a current state X, V in place
    of the scene's own (e.g. from ImplicitCloth.get_state())."""
    X = sc.X if X is None else X
    V = sc.V if V is None else V
    meta = {"name": sc.name, "nx": sc.nx, "ny": sc.ny, "steps": sc.steps,
            "params": dataclasses.asdict(sc.params)}
    np.savez(path, X=np.asarray(X, np.float32), V=np.asarray(V, np.float32),
             springs=np.asarray(sc.springs, SPRING_DTYPE), pinned=np.asarray(sc.pinned, np.uint8),
             triangles=grid_triangles(sc.nx, sc.ny), meta=np.array(json.dumps(meta)))


def load_scene(path: str) -> Scene:
    """Inverse of save_scene (the saved state becomes the scene's X, V)."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        p = meta["params"]
        p["gravity"] = tuple(p["gravity"])
        return Scene(meta["name"], int(meta["nx"]), int(meta["ny"]), z["X"].copy(), z["V"].copy(),
                     z["springs"].copy(), z["pinned"].copy(), StepParams(**p), int(meta["steps"]))
