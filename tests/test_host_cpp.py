"""The C++ host layer (include/opencloth.hpp) and the demos built on it
(opencloth_amd/host/{ie,bw,vl,pbd,vg}_demo.cpp): each program's scene, made by its
own C++ restatement of the reference demo's InitGL, equals the Python demo
scene bit for bit, so the GPU tests of their StepPhysics loops run the
oracle on the same inputs (CPU only: --scene-only makes no GPU call)."""
import struct
import subprocess

import numpy as np

from opencloth_amd import scenes
from opencloth_amd.build import host_program


def read_scene(path):
    b = open(path, "rb").read()
    n, s = struct.unpack("<ii", b[:8])
    off = 8
    X = np.frombuffer(b, np.float32, 3 * n, off).reshape(n, 3); off += 12 * n
    springs = np.frombuffer(b, scenes.SPRING_DTYPE, s, off); off += 24 * s
    pinned = np.frombuffer(b, np.uint8, n, off); off += n
    ell = np.frombuffer(b, np.float32, 36, off); off += 144
    assert off == len(b)
    return X, springs, pinned, ell


def test_demo_scene_matches(tmp_path):
    out = tmp_path / "scene.bin"
    r = subprocess.run([host_program("ie_demo"), "--scene-only", "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    X, springs, pinned, ell = read_scene(out)
    sc = scenes.demo_scene()
    e = sc.extra["ellipsoid"]
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(springs.view(np.uint8), sc.springs.view(np.uint8))
    assert np.array_equal(pinned, sc.pinned)
    assert np.array_equal(ell[:16].view(np.uint32), e["M"].view(np.uint32))
    assert np.array_equal(ell[16:32].view(np.uint32), e["Minv"].view(np.uint32))
    assert np.array_equal(ell[32:35], e["center"]) and ell[35] == e["radius"]


def test_bw_demo_scene_matches(tmp_path):
    out = tmp_path / "scene.bin"
    r = subprocess.run([host_program("bw_demo"), "--scene-only", "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    X, springs, pinned, ell = read_scene(out)
    sc = scenes.bw_demo_scene()
    e = sc.extra["ellipsoid"]
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(springs.view(np.uint8), sc.springs.view(np.uint8))
    assert np.array_equal(pinned, sc.pinned)
    assert np.array_equal(ell[:16].view(np.uint32), e["M"].view(np.uint32))


def test_verlet_demo_scene_matches(tmp_path):
    out = tmp_path / "scene.bin"
    r = subprocess.run([host_program("vl_demo"), "--scene-only", "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    X, springs, pinned, ell = read_scene(out)
    sc = scenes.verlet_demo_scene()
    e = sc.extra["ellipsoid"]
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(springs.view(np.uint8), sc.springs.view(np.uint8))
    assert np.array_equal(pinned, sc.pinned)
    assert np.array_equal(ell[:16].view(np.uint32), e["M"].view(np.uint32))
    assert np.array_equal(ell[16:32].view(np.uint32), e["Minv"].view(np.uint32))


def test_grid_verlet_demo_scene_matches(tmp_path):
    out = tmp_path / "scene.bin"
    r = subprocess.run([host_program("vg_demo"), "--scene-only", "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    b = open(out, "rb").read()
    n, u, v = struct.unpack("<iii", b[:12])
    X = np.frombuffer(b, np.float32, 4 * n, 12).reshape(n, 4)
    prm = np.frombuffer(b, np.float32, 14, 12 + 16 * n)
    assert 12 + 16 * n + 56 == len(b)
    sc = scenes.vgrid_demo_scene(64)
    assert (n, u, v) == (sc.width * sc.height, sc.width, sc.height)
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    p = sc.params
    expect = np.array([p["mass"], *p["gravity"], p["damping"], p["dt"], *p["spacing"], *p["ks"], *p["kd"]],
                      np.float32)
    assert np.array_equal(prm.view(np.uint32), expect.view(np.uint32))


def test_pbd_demo_scene_matches(tmp_path):
    out = tmp_path / "scene.bin"
    r = subprocess.run([host_program("pbd_demo"), "--scene-only", "--out", str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    b = open(out, "rb").read()
    n, nd, nb = struct.unpack("<iii", b[:12])
    off = 12
    X = np.frombuffer(b, np.float32, 3 * n, off).reshape(n, 3); off += 12 * n
    d = np.frombuffer(b, scenes.PBD_DISTANCE_DTYPE, nd, off); off += 12 * nd
    bend = np.frombuffer(b, scenes.PBD_BENDING_DTYPE, nb, off); off += 16 * nb
    pinned = np.frombuffer(b, np.uint8, n, off); off += n
    ell = np.frombuffer(b, np.float32, 36, off); off += 144
    assert off == len(b)
    sc = scenes.pbd_demo_scene()
    assert np.array_equal(X.view(np.uint32), sc.X.view(np.uint32))
    assert np.array_equal(d.view(np.uint8), sc.distance.view(np.uint8))
    assert np.array_equal(bend.view(np.uint8), sc.bending.view(np.uint8))
    assert np.array_equal(pinned, sc.pinned)
    assert np.array_equal(ell[16:32].view(np.uint32), sc.extra["ellipsoid"]["Minv"].view(np.uint32))


def test_demo_rejects_bad_arguments(tmp_path):
    r = subprocess.run([host_program("ie_demo"), "--steps"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_every_class_instantiates(tmp_path):
    """All four classes and every member compile when instantiated and
    link against the library (host compile and link only; nothing runs)."""
    import os

    from opencloth_amd.build import LIBDIR

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "all.cpp"
    src.write_text('''#include "opencloth.hpp"
struct v3 { float x, y, z; };
template class opencloth::ImplicitEuler<v3>;
template class opencloth::Verlet<v3>;
template class opencloth::PositionBasedDynamics<v3>;
struct v4 { float x, y, z, w; };
template class opencloth::GridVerlet<v4>;
int main(int argc, char**) {
    if (argc > 1) {  // never taken: only checks that every call links
        std::vector<v3> X(4), V(4), F;
        std::vector<oc_spring> s;
        std::vector<uint8_t> pin(4, 0);
        oc_params p = {};
        opencloth::ImplicitEuler<v3> a(1, 1, X, V, s, pin, p);
        a.ComputeForces(F); a.IntegrateImplicit(0.01f); a.AssembleSystem(0.01f);
        a.SolveConjugateGradient(); a.UpdateState(0.01f); a.ApplyProvotDynamicInverse();
        oc_ellipsoid e = {}; a.EllipsoidCollision(e); a.StepPhysics(0.01f, &e); a.Step(2); a.Upload();
        int it, st; a.LastSolve(&it, &st);
        oc_verlet_params vp = {};
        opencloth::Verlet<v3> b(1, 1, X, V, s, pin, vp);
        b.ComputeForces(F); b.StepPhysics(0.01f, &e); b.Upload();
        oc_pbd_params pp = {};
        std::vector<oc_pbd_distance> d; std::vector<oc_pbd_bending> bd;
        opencloth::PositionBasedDynamics<v3> c(X, V, pin, d, bd, pp);
        c.StepPhysics(0.01f, &e); c.Upload();
        std::vector<v4> X4(4), XL4(4);
        oc_vgrid_params gp = {};
        opencloth::GridVerlet<v4> g(2, 2, X4, XL4, gp);
        g.Verlet(nullptr); g.Step(3); g.SetParams(gp); g.Upload(); g.Download();
    }
    return 0;
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(tmp_path / "all"),
                        "-L", LIBDIR, "-lopencloth", "-Wl,-rpath-link,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_compiles_with_a_glm_like_type(tmp_path):
    """ImplicitEuler<Vec3> accepts any 12-byte three-float type and rejects
    others at compile time (host compile only)."""
    import os

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    good = tmp_path / "good.cpp"
    good.write_text('#include "opencloth.hpp"\nstruct v3 { float x, y, z; };\n'
                    "float* f(std::vector<v3>& a) { return opencloth::floats(a); }\n")
    bad = tmp_path / "bad.cpp"
    bad.write_text('#include "opencloth.hpp"\nstruct v4 { float x, y, z, w; };\n'
                   "float* f(std::vector<v4>& a) { return opencloth::floats(a); }\n")
    ok = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, str(good)],
                        capture_output=True, text=True)
    assert ok.returncode == 0, ok.stderr
    no = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, str(bad)],
                        capture_output=True, text=True)
    assert no.returncode != 0 and "three packed floats" in no.stderr
