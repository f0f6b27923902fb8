"""ctypes front end of the CPU oracle (oracle.c, pbd.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the opencloth_amd package.

`Oracle(32)` is the FP32 restatement, `Oracle(64)` the FP64 oracle; both take
float32 scene arrays and compute in their own precision.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF = os.path.join(HERE, "_ref")

NOT_CONVERGED, BREAKDOWN, NONFINITE = 1, 2, 4
JACOBI = 4  # Params.flags bit: Jacobi-preconditioned CG
CG1 = 16    # Params.flags bit: single-reduction CG (Chronopoulos-Gear)


class Params(C.Structure):
    _fields_ = [
        ("mass", C.c_double), ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
        ("h", C.c_double), ("tol", C.c_double),
        ("max_it", C.c_int32), ("tol_mode", C.c_int32), ("warm_start", C.c_int32),
        ("dot_mode", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("tx", C.c_int32),
        ("ty", C.c_int32), ("flags", C.c_int32), ("pad", C.c_int32), ("damping", C.c_double),
    ]


def build(force: bool = False) -> None:
    """Compile the oracle (and the GLM harness when the reference is mounted)."""
    libs = [os.path.join(BUILD, "liboracle32.so"), os.path.join(BUILD, "liboracle64.so")]
    if force or not all(os.path.exists(p) for p in libs):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref() -> bool:
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return os.path.exists(os.path.join(REF, "libglmref.so"))


_P = C.c_void_p


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, bits: int = 32):
        assert bits in (32, 64)
        self.bits = bits
        self.dtype = np.float32 if bits == 32 else np.float64
        path = os.path.join(BUILD, f"liboracle{bits}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        L.orc_create.restype = _P
        L.orc_create.argtypes = [C.c_int, C.c_int, _P, _P]
        L.orc_destroy.argtypes = [_P]
        L.orc_num_pairs.argtypes = [_P]
        L.orc_pairs.argtypes = [_P, _P, _P, _P]
        L.orc_grid_positions.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, _P]
        L.orc_grid_springs.argtypes = [C.c_int, C.c_int, _P] + [C.c_float] * 6 + [_P]
        L.orc_spring_terms.argtypes = [_P, _P, _P, _P, _P, _P]
        L.orc_forces_scatter.argtypes = [_P, _P, _P, C.c_double, _P, C.c_double, _P]
        L.orc_forces_gather.argtypes = [_P, _P, _P, C.c_double, _P, C.c_double, _P, _P]
        L.orc_assemble.argtypes = [_P, _P, _P, _P, _P, _P, _P]
        L.orc_matvec.argtypes = [_P, _P, C.c_double, _P, _P]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [_P, _P, _P, _P]
        L.orc_cg.argtypes = [_P, _P, _P, _P, _P, _P, _P]
        L.orc_cg_blockdiag.argtypes = [C.c_int, _P, _P, _P, C.c_int, C.c_double]
        L.orc_cg_reference_blockdiag.argtypes = [C.c_int, _P, _P, _P, C.c_int, C.c_double]
        L.orc_step.argtypes = [_P, _P, _P, _P, _P, _P]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_ie_literal_step.argtypes = [_P, _P, _P, _P, C.c_double, _P, C.c_double, C.c_int, C.c_double]
        L.orc_rollout.argtypes = [_P, _P, _P, _P, _P, C.c_int, _P, _P]
        L.orc_set_ks.argtypes = [_P, _P]
        L.orc_ellipsoid.argtypes = [_P, _P, _P, _P, _P, _P, C.c_float, C.c_int]
        L.orc_provot.argtypes = [_P, _P, _P, C.c_double]
        L.orc_verlet_forces.argtypes = [_P, _P, _P, C.c_double, C.c_double, _P, C.c_double, _P]
        L.orc_verlet_integrate.argtypes = [_P, _P, _P, _P, C.c_double, C.c_double, C.c_int]
        L.orc_verlet_step.argtypes = [_P, _P, _P, C.c_double, C.c_double, _P, C.c_double, C.c_int,
                                      _P, _P, _P, C.c_float]
        L.orc_step_adjoint.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_ellipsoid_points.argtypes = [C.c_int, _P, _P, _P, _P, _P, C.c_float, C.c_int]
        L.orc_pbd_create.restype = _P
        L.orc_pbd_create.argtypes = [C.c_int, _P, _P, C.c_double, _P, C.c_int, _P, C.c_int, C.c_int]
        L.orc_pbd_destroy.argtypes = [_P]
        L.orc_pbd_colours.argtypes = [_P, _P, _P, _P, _P, _P, _P]
        L.orc_pbd_rest.argtypes = [_P, _P, _P, _P, _P, _P]
        L.orc_pbd_predict.argtypes = [_P, _P, _P, _P, _P, C.c_int, _P]
        L.orc_pbd_project.argtypes = [_P, _P, _P, C.c_int]
        L.orc_pbd_integrate.argtypes = [_P, _P, _P, _P, _P]
        L.orc_pbd_step.argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, _P, _P, _P, C.c_float]
        L.orc_vgrid_step.argtypes = [C.c_int, C.c_int, _P, _P, _P, _P, C.c_int]

    def set_threads(self, n: int) -> int:
        """Worker threads of the implicit step's parallel loops (results are
        bitwise the same for any count); returns the previous count."""
        return int(self.L.orc_set_threads(int(n)))

    # -- arrays --------------------------------------------------------------
    def arr(self, a) -> np.ndarray:
        return np.ascontiguousarray(a, dtype=self.dtype)

    # -- scene helpers ---------------------------------------------------------
    def grid_positions(self, nx, ny, fullsize=4.0, y0=0.0):
        X = np.zeros((nx * ny, 3), np.float32)
        self.L.orc_grid_positions(nx, ny, fullsize, y0, _ptr(X))
        return X

    def grid_springs(self, nx, ny, X, ks=(50.0, 50.0, 50.0), kd=(-0.25, -0.25, -0.25)):
        from opencloth_amd.scenes import SPRING_DTYPE

        X = np.ascontiguousarray(X, np.float32)
        n = self.L.orc_grid_springs(nx, ny, _ptr(X), ks[0], kd[0], ks[1], kd[1], ks[2], kd[2], None)
        out = np.zeros(n, SPRING_DTYPE)
        self.L.orc_grid_springs(nx, ny, _ptr(X), ks[0], kd[0], ks[1], kd[1], ks[2], kd[2], _ptr(out))
        return out

    # -- systems ---------------------------------------------------------------
    def system(self, n_vertices, springs, pinned=None) -> "System":
        return System(self, n_vertices, springs, pinned)

    def pbd(self, X0, pinned, mass, distance, bending, solver_iterations) -> "PbdSystem":
        return PbdSystem(self, X0, pinned, mass, distance, bending, solver_iterations)

    def vgrid_rollout(self, width, height, X4, XL4, pinned, prm, nsteps):
        """nsteps of the GPU demos' Verlet kernel (orc_vgrid_step) on copies of
        the (N, 4) position arrays; prm: dict with mass, gravity, damping, dt,
        spacing, ks, kd.  Returns (X4, XL4) in this oracle's dtype."""
        X = self.arr(X4).copy()
        XL = self.arr(XL4).copy()
        pin = np.ascontiguousarray(pinned, np.uint8)
        assert X.shape == XL.shape == (width * height, 4) and pin.shape == (width * height,)
        p = np.array([prm["mass"], *prm["gravity"], prm["damping"], prm["dt"], *prm["spacing"],
                      *prm["ks"], *prm["kd"]], np.float64)
        self.L.orc_vgrid_step(int(width), int(height), _ptr(X), _ptr(XL), _ptr(pin), _ptr(p),
                              int(nsteps))
        return X, XL


def make_params(p, nx=0, ny=0, tx=0, ty=0, dot_mode=0, **over) -> Params:
    """Params from an opencloth_amd.scenes.StepParams (or keyword overrides)."""
    q = Params()
    q.mass = float(over.get("mass", p.mass))
    g = over.get("gravity", p.gravity)
    q.gx, q.gy, q.gz = float(g[0]), float(g[1]), float(g[2])
    q.h = float(over.get("dt", p.dt))
    q.tol = float(over.get("cg_tol", p.cg_tol))
    q.max_it = int(over.get("cg_max_iters", p.cg_max_iters))
    q.tol_mode = int(over.get("cg_tol_mode", p.cg_tol_mode))
    q.warm_start = int(bool(over.get("warm_start", p.warm_start)))
    q.dot_mode = dot_mode
    q.nx, q.ny, q.tx, q.ty = nx, ny, tx, ty
    q.flags = (1 if over.get("ground_clamp", p.ground_clamp) else 0) | (
        JACOBI if over.get("cg_jacobi", getattr(p, "cg_jacobi", False)) else 0) | (
        CG1 if over.get("cg_single_reduction", getattr(p, "cg_single_reduction", False)) else 0)
    q.damping = float(over.get("damping", getattr(p, "damping", 0.0)))
    return q


class System:
    """One cloth's topology inside the oracle (orc_sys)."""

    def __init__(self, o: Oracle, n_vertices, springs, pinned=None):
        self.o = o
        self.N = int(n_vertices)
        self.springs = np.ascontiguousarray(springs)
        self.S = len(springs)
        self.pinned = (np.zeros(self.N, np.uint8) if pinned is None
                       else np.ascontiguousarray(pinned, np.uint8))
        self.h = o.L.orc_create(self.N, self.S, _ptr(self.springs), _ptr(self.pinned))
        self.P = o.L.orc_num_pairs(self.h)

    def __del__(self):
        try:
            self.o.L.orc_destroy(self.h)
        except Exception:
            pass

    def pairs(self):
        a = np.zeros(self.P, np.int32); b = np.zeros(self.P, np.int32)
        sp = np.zeros(self.S, np.int32)
        self.o.L.orc_pairs(self.h, _ptr(a), _ptr(b), _ptr(sp))
        return a, b, sp

    def spring_terms(self, X, V):
        o = self.o
        X = o.arr(X); V = o.arr(V)
        F = np.zeros((self.S, 3), o.dtype); K = np.zeros((self.S, 6), o.dtype)
        P = np.zeros((self.S, 6), o.dtype)
        o.L.orc_spring_terms(self.h, _ptr(X), _ptr(V), _ptr(F), _ptr(K), _ptr(P))
        return F, K, P

    def forces_scatter(self, X, V, mass, g, damping=0.0):
        o = self.o
        X = o.arr(X); V = o.arr(V)
        F = np.zeros((self.N, 3), o.dtype)
        gg = np.asarray(g, np.float64)
        o.L.orc_forces_scatter(self.h, _ptr(X), _ptr(V), float(mass), _ptr(gg), float(damping), _ptr(F))
        return F

    def forces_gather(self, X, V, mass, g, damping=0.0):
        o = self.o
        X = o.arr(X); V = o.arr(V)
        F = np.zeros((self.N, 3), o.dtype); KV = np.zeros((self.N, 3), o.dtype)
        gg = np.asarray(g, np.float64)
        o.L.orc_forces_gather(self.h, _ptr(X), _ptr(V), float(mass), _ptr(gg), float(damping), _ptr(F),
                              _ptr(KV))
        return F, KV

    def assemble(self, X, V, prm: Params, with_diag=False):
        o = self.o
        X = o.arr(X); V = o.arr(V)
        B = np.zeros((self.N, 3), o.dtype); Cp = np.zeros((self.P, 6), o.dtype)
        D = np.zeros((self.N, 3), o.dtype) if with_diag else None
        o.L.orc_assemble(self.h, _ptr(X), _ptr(V), C.byref(prm), _ptr(B), _ptr(Cp),
                         _ptr(D) if D is not None else None)
        return (B, Cp, D) if with_diag else (B, Cp)

    def matvec(self, Cp, mass, d):
        o = self.o
        d = o.arr(d); Cp = o.arr(Cp)
        q = np.zeros((self.N, 3), o.dtype)
        o.L.orc_matvec(self.h, _ptr(Cp), float(mass), _ptr(d), _ptr(q))
        return q

    def dot(self, prm: Params, a, b) -> float:
        o = self.o
        a = o.arr(a); b = o.arr(b)
        return o.L.orc_dot(self.h, C.byref(prm), _ptr(a), _ptr(b))

    def cg(self, Cp, prm: Params, b, dv0=None):
        o = self.o
        Cp = o.arr(Cp); b = o.arr(b)
        dv = np.zeros((self.N, 3), o.dtype) if dv0 is None else o.arr(dv0).copy()
        st = C.c_int(0); dl = C.c_double(0)
        it = o.L.orc_cg(self.h, _ptr(Cp), C.byref(prm), _ptr(b), _ptr(dv), C.byref(st), C.byref(dl))
        return dv, it, st.value, dl.value

    def step(self, X, V, DV, prm: Params):
        """One step in place on (X, V, DV) arrays of this oracle's dtype."""
        st = C.c_int(0)
        it = self.o.L.orc_step(self.h, _ptr(X), _ptr(V), _ptr(DV), C.byref(prm), C.byref(st))
        return it, st.value

    def rollout(self, X, V, prm: Params, nsteps, DV=None, history=False):
        """Run nsteps from (X, V); returns (X, V, DV, hist or None, iters, status)."""
        o = self.o
        X = o.arr(X).copy(); V = o.arr(V).copy()
        DV = np.zeros((self.N, 3), o.dtype) if DV is None else o.arr(DV).copy()
        hist = np.zeros((nsteps, self.N, 3), o.dtype) if history else None
        iters = np.zeros(nsteps, np.int32)
        st = o.L.orc_rollout(self.h, _ptr(X), _ptr(V), _ptr(DV), C.byref(prm), nsteps,
                             _ptr(hist) if hist is not None else None, _ptr(iters))
        return X, V, DV, hist, iters, st


    def ie_literal_step(self, X, V, Vnew, mass, gravity, h, i_max=10, eps=1e-3):
        """The reference's own ComputeForces + IntegrateImplicit, bug for bug
        (spring-indexed blocks), in place on X, V and the warm start Vnew;
        returns the CG iterations.  For the record and the CPU timing."""
        g = np.ascontiguousarray(gravity, np.float64)
        it = self.o.L.orc_ie_literal_step(self.h, _ptr(X), _ptr(V), _ptr(Vnew), float(mass), _ptr(g),
                                          float(h), int(i_max), float(eps))
        if it < 0:
            raise ValueError("the literal step needs at least as many springs as vertices")
        return it

    def set_ks(self, ks):
        """Per-spring stiffness in the oracle's precision."""
        k = np.ascontiguousarray(ks, np.float64)
        self.o.L.orc_set_ks(self.h, _ptr(k))

    def step_adjoint(self, X, V, DV, prm: Params, gx_out, gv_out, gks):
        """Adjoint of one step at pre-step state (X, V) with solution DV.
        Returns (gx_in, gv_in, mu, iters); gks (S,) is accumulated in place."""
        o = self.o
        X = o.arr(X); V = o.arr(V); DV = o.arr(DV)
        gxo = o.arr(gx_out); gvo = o.arr(gv_out)
        gx = np.zeros((self.N, 3), o.dtype); gv = np.zeros((self.N, 3), o.dtype)
        mu = np.zeros((self.N, 3), o.dtype)
        st = C.c_int(0)
        it = o.L.orc_step_adjoint(self.h, _ptr(X), _ptr(V), _ptr(DV), C.byref(prm), _ptr(gxo),
                                  _ptr(gvo), _ptr(gx), _ptr(gv), _ptr(gks), _ptr(mu), C.byref(st))
        return gx, gv, mu, it

    # -- post-step passes and Verlet -------------------------------------------
    def ellipsoid(self, X, W, ell, mode):
        """In place on X, W (this oracle's dtype).  ell: dict with M, Minv
        (column-major 16), center (3), radius."""
        M = np.ascontiguousarray(ell["M"], np.float32); Mi = np.ascontiguousarray(ell["Minv"], np.float32)
        cen = np.ascontiguousarray(ell["center"], np.float32)
        self.o.L.orc_ellipsoid(self.h, _ptr(X), _ptr(W), _ptr(M), _ptr(Mi), _ptr(cen),
                               float(ell["radius"]), int(mode))

    def provot(self, X, V, slack=1.0):
        self.o.L.orc_provot(self.h, _ptr(X), _ptr(V), float(slack))

    def verlet_forces(self, X, XL, dt, mass, g, damping):
        o = self.o
        X = o.arr(X); XL = o.arr(XL)
        F = np.zeros((self.N, 3), o.dtype)
        gg = np.asarray(g, np.float64)
        o.L.orc_verlet_forces(self.h, _ptr(X), _ptr(XL), float(dt), float(mass), _ptr(gg),
                              float(damping), _ptr(F))
        return F

    def verlet_rollout(self, X, XL, dt, mass, g, damping, flags, nsteps, ell=None):
        o = self.o
        X = o.arr(X).copy(); XL = o.arr(XL).copy()
        gg = np.asarray(g, np.float64)
        if ell is not None:
            M = np.ascontiguousarray(ell["M"], np.float32); Mi = np.ascontiguousarray(ell["Minv"], np.float32)
            cen = np.ascontiguousarray(ell["center"], np.float32); rad = float(ell["radius"])
        for _ in range(nsteps):
            if ell is None:
                o.L.orc_verlet_step(self.h, _ptr(X), _ptr(XL), float(dt), float(mass), _ptr(gg),
                                    float(damping), int(flags), None, None, None, 0.0)
            else:
                o.L.orc_verlet_step(self.h, _ptr(X), _ptr(XL), float(dt), float(mass), _ptr(gg),
                                    float(damping), int(flags), _ptr(M), _ptr(Mi), _ptr(cen), rad)
        return X, XL

    def step_physics(self, X, V, DV, prm: Params, ell=None, provot_slack=0.0):
        """StepPhysics (IE:774-779): implicit step, ellipsoid (zero velocity),
        Provot (when slack > 0).  In place; returns (iters, status)."""
        it, st = self.step(X, V, DV, prm)
        if ell is not None:
            self.ellipsoid(X, V, ell, 0)
        if provot_slack > 0:
            self.provot(X, V, provot_slack)
        return it, st

    def rollout_forward(self, X0, V0, prm: Params, nsteps):
        """Rollout keeping every pre-step state and solution: (X_T, V_T, states)."""
        o = self.o
        X = o.arr(X0).copy(); V = o.arr(V0).copy(); DV = np.zeros((self.N, 3), o.dtype)
        states = []
        for _ in range(nsteps):
            Xn, Vn = X.copy(), V.copy()
            self.step(X, V, DV, prm)
            states.append((Xn, Vn, DV.copy()))
        return X, V, states

    def rollout_grad(self, X0, V0, prm: Params, nsteps, gx_T, gv_T=None, states=None):
        """Gradients of a scalar L through nsteps given dL/dx_T (and dL/dv_T):
        returns (dL/dks, dL/dx0, dL/dv0)."""
        o = self.o
        if states is None:
            _, _, states = self.rollout_forward(X0, V0, prm, nsteps)
        gks = np.zeros(self.S, o.dtype)
        gx = o.arr(gx_T).copy()
        gv = np.zeros((self.N, 3), o.dtype) if gv_T is None else o.arr(gv_T).copy()
        for Xn, Vn, DVn in reversed(states):
            gx, gv, _, _ = self.step_adjoint(Xn, Vn, DVn, prm, gx, gv, gks)
        return gks, gx, gv


class PbdParams(C.Structure):
    _fields_ = [
        ("mass", C.c_double), ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
        ("dt", C.c_double), ("global_damping", C.c_double), ("k_damp", C.c_double),
        ("solver_iterations", C.c_int32), ("ground", C.c_int32),
    ]


def make_pbd_params(p, **over) -> PbdParams:
    """From opencloth_amd.scenes.PBDParams (or anything with its fields)."""
    q = PbdParams()
    q.mass = p.mass
    q.gx, q.gy, q.gz = p.gravity
    q.dt = p.dt
    q.global_damping = p.global_damping
    q.k_damp = p.k_damp
    q.solver_iterations = p.solver_iterations
    q.ground = 1 if p.ground else 0
    for k, v in over.items():
        setattr(q, k, v)
    return q


class PbdSystem:
    """The PBD step of pbd.c.  order 0: the reference's sequential order;
    order 1: the HIP kernels' colour order and reduction tree."""

    def __init__(self, o: Oracle, X0, pinned, mass, distance, bending, solver_iterations):
        self.o = o
        X0 = np.ascontiguousarray(X0, np.float32)
        self.N = len(X0)
        self.pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        self.dist = np.ascontiguousarray(distance)
        self.bend = np.ascontiguousarray(bending)
        self.nd, self.nb = len(self.dist), len(self.bend)
        self.h = o.L.orc_pbd_create(self.N, _ptr(X0), None if self.pins is None else _ptr(self.pins),
                                    float(mass), _ptr(self.dist), self.nd, _ptr(self.bend), self.nb,
                                    int(solver_iterations))
        if not self.h:
            raise ValueError("orc_pbd_create rejected the constraints")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.o.L.orc_pbd_destroy(h)
            self.h = None

    def colours(self):
        """(distance order, distance colour offsets, bending order, bending offsets)."""
        nd, nb = C.c_int(), C.c_int()
        self.o.L.orc_pbd_colours(self.h, C.byref(nd), C.byref(nb), None, None, None, None)
        do = np.zeros(self.nd, np.int32); dp = np.zeros(nd.value + 1, np.int32)
        bo = np.zeros(self.nb, np.int32); bp = np.zeros(nb.value + 1, np.int32)
        self.o.L.orc_pbd_colours(self.h, C.byref(nd), C.byref(nb), _ptr(do), _ptr(dp), _ptr(bo), _ptr(bp))
        return do, dp, bo, bp

    def rest(self):
        dt = self.o.dtype
        d_rest = np.zeros(self.nd, dt); d_kp = np.zeros(self.nd, dt)
        b_rest = np.zeros(self.nb, dt); b_w = np.zeros(self.nb, dt); b_kp = np.zeros(self.nb, dt)
        self.o.L.orc_pbd_rest(self.h, _ptr(d_rest), _ptr(d_kp), _ptr(b_rest), _ptr(b_w), _ptr(b_kp))
        return d_rest, d_kp, b_rest, b_w, b_kp

    def predict(self, X, V, prm: PbdParams, order):
        """Returns (V after damping, predicted positions T, angular velocity w)."""
        o = self.o
        X = o.arr(X); V = o.arr(V).copy()
        T = np.zeros_like(V); w = np.zeros(3, o.dtype)
        o.L.orc_pbd_predict(self.h, _ptr(X), _ptr(V), _ptr(T), C.byref(prm), int(order), _ptr(w))
        return V, T, w

    def project(self, T, prm: PbdParams, order):
        T = self.o.arr(T).copy()
        self.o.L.orc_pbd_project(self.h, _ptr(T), C.byref(prm), int(order))
        return T

    def step(self, X, V, prm: PbdParams, order, nsteps=1, ell=None):
        o = self.o
        X = o.arr(X).copy(); V = o.arr(V).copy()
        if ell is None:
            o.L.orc_pbd_step(self.h, _ptr(X), _ptr(V), C.byref(prm), int(order), int(nsteps),
                             None, None, None, 0.0)
        else:
            M = np.ascontiguousarray(ell["M"], np.float32); Mi = np.ascontiguousarray(ell["Minv"], np.float32)
            cen = np.ascontiguousarray(ell["center"], np.float32)
            o.L.orc_pbd_step(self.h, _ptr(X), _ptr(V), C.byref(prm), int(order), int(nsteps),
                             _ptr(M), _ptr(Mi), _ptr(cen), float(ell["radius"]))
        return X, V


def cg_blockdiag(o: Oracle, A, b, x0, i_max, eps, literal=False):
    A = o.arr(A); b = o.arr(b); x = o.arr(x0).copy()
    f = o.L.orc_cg_reference_blockdiag if literal else o.L.orc_cg_blockdiag
    it = f(len(b), _ptr(A), _ptr(b), _ptr(x), i_max, eps)
    return x, it


class GlmRef:
    """The GLM 0.9.0 harness built from the reference's headers (oracle/_ref)."""

    def __init__(self):
        path = os.path.join(REF, "libglmref.so")
        self.L = C.CDLL(path)
        self.L.glmref_ops.argtypes = [C.c_int] + [_P] * 8
        self.L.glmref_spring.argtypes = [C.c_int] + [_P] * 9

    @staticmethod
    def available() -> bool:
        return os.path.exists(os.path.join(REF, "libglmref.so"))

    def ops(self, a, b, M):
        n = len(a)
        a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
        M = np.ascontiguousarray(M, np.float32)
        d = np.zeros(n, np.float32); l = np.zeros(n, np.float32)
        nn = np.zeros((n, 3), np.float32); o = np.zeros((n, 9), np.float32)
        mv = np.zeros((n, 3), np.float32)
        self.L.glmref_ops(n, _ptr(a), _ptr(b), _ptr(M), _ptr(d), _ptr(l), _ptr(nn), _ptr(o), _ptr(mv))
        return d, l, nn, o, mv

    def spring(self, x1, x2, v1, v2, L0, Ks, Kd):
        n = len(x1)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x1, x2, v1, v2, L0, Ks, Kd)]
        f = np.zeros((n, 3), np.float32); K = np.zeros((n, 9), np.float32)
        self.L.glmref_spring(n, *[_ptr(a) for a in arrs], _ptr(f), _ptr(K))
        return f, K
