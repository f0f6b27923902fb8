"""The batched stencil assembly's slot rule (CPU only): on the reference's
grid spring lists, a vertex's springs sit in the assembly slots that the
mask's popcount gives, so batch.hip can read their constants without the
slot tables.  A small C++ program over the library's topology code checks
every vertex; a shuffled spring list, which has no stencil form, is
reported as such."""
import os
import subprocess

import numpy as np
import pytest

from opencloth_amd import scenes
from opencloth_amd.build import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("slots") / "stencil_slots")
    csrc = os.path.join(ROOT, "opencloth_amd", "csrc")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", csrc,
           os.path.join(HERE, "cpp", "stencil_slots.cpp"), os.path.join(csrc, "topology.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    return out


def run(exe, tmp_path, nx, ny, springs):
    path = str(tmp_path / f"springs_{nx}x{ny}.bin")
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, len(springs)], np.int32).tobytes())
        f.write(springs.tobytes())
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.strip()


@pytest.mark.parametrize("nx,ny", [(3, 3), (3, 5), (16, 16), (21, 21), (37, 90), (128, 17)])
def test_slot_rule_on_grid_springs(exe, tmp_path, nx, ny):
    X = scenes.grid_positions(nx, ny)
    assert run(exe, tmp_path, nx, ny, scenes.grid_springs(nx, ny, X)) == "ok"


def test_shuffled_list_has_no_stencil_form(exe, tmp_path):
    X = scenes.grid_positions(21, 21)
    sp = scenes.grid_springs(21, 21, X)
    sp = sp[np.random.default_rng(3).permutation(len(sp))]
    assert run(exe, tmp_path, 21, 21, sp) in ("ok", "no stencil assembly form")


@pytest.fixture(scope="module")
def ring_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ring") / "band_ring")
    csrc = os.path.join(ROOT, "opencloth_amd", "csrc")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", csrc,
           os.path.join(HERE, "cpp", "band_ring.cpp"), os.path.join(csrc, "topology.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    return out


@pytest.mark.parametrize("nx,ny,T,NT", [(16, 16, 256, 512), (16, 16, 64, 512), (21, 21, 64, 256),
                                        (37, 90, 512, 512), (128, 128, 512, 512), (128, 17, 64, 512),
                                        (100, 33, 1024, 1024), (3, 5, 64, 1024)])
def test_band_rings(ring_exe, tmp_path, nx, ny, T, NT):
    """The batched spring-once assembly's bands and rings (CPU walk-through):
    every partner reads its own owner's result, inside the LDS allocation,
    also where T is shorter than a ring's offset (two grid rows)."""
    X = scenes.grid_positions(nx, ny)
    sp = scenes.grid_springs(nx, ny, X)
    path = str(tmp_path / "springs.bin")
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, len(sp)], np.int32).tobytes())
        f.write(sp.tobytes())
    r = subprocess.run([ring_exe, path, str(T), str(NT)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
