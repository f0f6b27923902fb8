"""Python mirror of the reference's implicit-step interface over the C ABI.

The reference drives one cloth held in globals through free functions
(OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp):
ComputeForces (:645), IntegrateImplicit (:684), SolveConjugateGradient
(:227), StepPhysics (:774).  `ImplicitCloth` keeps the same operations on a
handle whose state lives in device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import OcEllipsoid, OcParams, OcPbdParams, OcVerletParams, OcVgridParams, check, lib, ptr
from .scenes import (PBD_BENDING_DTYPE, PBD_DISTANCE_DTYPE, SPRING_DTYPE, PBDParams, Scene,
                     StepParams)


def to_oc_params(p: StepParams) -> OcParams:
    q = OcParams()
    q.mass = p.mass
    q.gravity[0], q.gravity[1], q.gravity[2] = p.gravity
    q.dt = p.dt
    q.cg_max_iters = p.cg_max_iters
    q.cg_tol = p.cg_tol
    q.cg_tol_mode = p.cg_tol_mode
    q.warm_start = 1 if p.warm_start else 0
    q.flags = ((_lib.OC_FLAG_GROUND_CLAMP if p.ground_clamp else 0)
               | (_lib.OC_FLAG_CG_PERSISTENT if p.cg_persistent else 0)
               | (_lib.OC_FLAG_JACOBI if p.cg_jacobi else 0)
               | (_lib.OC_FLAG_CG_SINGLE_REDUCTION if getattr(p, "cg_single_reduction", False) else 0))
    q.damping = getattr(p, "damping", 0.0)
    return q


def _vec(a, n):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != (n, 3):
        raise ValueError(f"expected shape ({n}, 3), got {a.shape}")
    return a


class ImplicitCloth:
    """One cloth on the current HIP device (oc_cloth)."""

    def __init__(self, nx, ny, X, springs, pinned=None, params: StepParams | None = None, V=None):
        self.nx, self.ny = int(nx), int(ny)
        self.N = self.nx * self.ny
        self.params = params or StepParams()
        X = _vec(X, self.N)
        V = None if V is None else _vec(V, self.N)
        springs = np.ascontiguousarray(springs, dtype=SPRING_DTYPE)
        pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        if pins is not None and pins.shape != (self.N,):
            raise ValueError("pinned must have one entry per vertex")
        self._p = to_oc_params(self.params)
        h = C.c_void_p()
        check(lib().oc_cloth_create(self.nx, self.ny, ptr(X), ptr(V), ptr(springs), len(springs),
                                    ptr(pins), C.byref(self._p), C.byref(h)))
        self.h = h
        self.S = len(springs)
        n = C.c_int32()
        check(lib().oc_cloth_num_pairs(self.h, C.byref(n)))
        self.P = n.value

    @classmethod
    def from_scene(cls, sc: Scene) -> "ImplicitCloth":
        return cls(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, sc.params, sc.V)

    def close(self):
        if getattr(self, "h", None):
            lib().oc_cloth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration -------------------------------------------------------
    def set_stream(self, stream_handle: int | None):
        check(lib().oc_cloth_set_stream(self.h, C.c_void_p(stream_handle or 0)))

    def set_params(self, params: StepParams):
        self.params = params
        self._p = to_oc_params(params)
        check(lib().oc_cloth_set_params(self.h, C.byref(self._p)))

    # -- state ---------------------------------------------------------------
    def set_state(self, X=None, V=None):
        Xa = None if X is None else _vec(X, self.N)
        Va = None if V is None else _vec(V, self.N)
        check(lib().oc_cloth_set_state(self.h, ptr(Xa), ptr(Va)))

    def get_state(self):
        X = np.zeros((self.N, 3), np.float32)
        V = np.zeros((self.N, 3), np.float32)
        check(lib().oc_cloth_get_state(self.h, ptr(X), ptr(V)))
        return X, V

    @property
    def X(self):
        return self.get_state()[0]

    @property
    def V(self):
        return self.get_state()[1]

    def get_dv(self):
        dv = np.zeros((self.N, 3), np.float32)
        check(lib().oc_cloth_get_dv(self.h, ptr(dv)))
        return dv

    def set_dv(self, dv=None):
        a = None if dv is None else _vec(dv, self.N)
        check(lib().oc_cloth_set_dv(self.h, ptr(a)))

    # -- the reference's operations ------------------------------------------
    def compute_forces(self) -> np.ndarray:
        """ComputeForces (IE:645-682): gravity plus spring forces, zero on pins."""
        F = np.zeros((self.N, 3), np.float32)
        check(lib().oc_cloth_compute_forces(self.h, ptr(F)))
        return F

    def assemble(self, dt=None):
        check(lib().oc_cloth_assemble(self.h, self.params.dt if dt is None else dt))

    def solve_conjugate_gradient(self):
        """SolveConjugateGradient (IE:227-249) on the assembled system;
        returns (iterations, status bits)."""
        it = C.c_int32(); st = C.c_int32()
        check(lib().oc_cloth_solve(self.h, C.byref(it), C.byref(st)))
        return it.value, st.value

    def update(self, dt=None):
        check(lib().oc_cloth_update(self.h, self.params.dt if dt is None else dt))

    def integrate_implicit(self, dt=None):
        """IntegrateImplicit (IE:684-700)."""
        check(lib().oc_cloth_integrate_implicit(self.h, self.params.dt if dt is None else dt))

    def step(self, nsteps: int = 1):
        """StepPhysics (IE:774-779) of the implicit integrator, nsteps times."""
        check(lib().oc_cloth_step(self.h, int(nsteps)))

    def synchronize(self):
        check(lib().oc_cloth_synchronize(self.h))

    KERNEL_CLASSES = ("assemble", "cg_spmv", "cg_axpy", "cg_finish", "update")

    def profile_steps(self, nsteps: int) -> dict:
        """nsteps steps exactly as step() runs them, each kernel dispatch
        timed by its own start/stop events: {kernel class: (ms, launches)}."""
        ms = (C.c_float * 5)(); n = (C.c_int32 * 5)()
        check(lib().oc_cloth_profile_steps(self.h, int(nsteps), ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def time_cg_spmv(self, reps: int = 200) -> float:
        """Average microseconds of one CG SpMV launch over reps back-to-back
        launches (oc_cloth_time_cg_spmv)."""
        t = C.c_float()
        check(lib().oc_cloth_time_cg_spmv(self.h, int(reps), C.byref(t)))
        return t.value

    KERNELS = ("assemble", "assemble_fused", "cg_spmv", "cg_axpy", "update")

    def time_kernel(self, kernel: str, reps: int = 200) -> float:
        """Average microseconds of one launch of a kernel class over reps
        back-to-back launches (oc_cloth_time_kernel)."""
        t = C.c_float()
        check(lib().oc_cloth_time_kernel(self.h, self.KERNELS.index(kernel), int(reps), C.byref(t)))
        return t.value

    def last_status(self):
        it = C.c_int32(); st = C.c_int32()
        check(lib().oc_cloth_last_status(self.h, C.byref(it), C.byref(st)))
        return it.value, st.value

    def sticky_status(self) -> int:
        """OR of the status bits of every solve and update since the last
        call (then cleared)."""
        st = C.c_int32()
        check(lib().oc_cloth_sticky_status(self.h, C.byref(st)))
        return st.value

    # -- post-step passes (StepPhysics, IE:774-779) ---------------------------
    def ellipsoid_collision(self, ell: dict):
        """EllipsoidCollision (IE:727-751)."""
        e = OcEllipsoid.from_dict(ell)
        check(lib().oc_cloth_ellipsoid_collision(self.h, C.byref(e)))

    def provot(self, slack: float = 1.0):
        """ApplyProvotDynamicInverse (IE:702-726; slack 1.01 for BW:808)."""
        check(lib().oc_cloth_provot(self.h, float(slack)))

    def step_physics(self, nsteps: int = 1, ell: dict | None = None, provot_slack: float = 0.0):
        """StepPhysics (IE:774-779): implicit step, ellipsoid, Provot."""
        e = None if ell is None else OcEllipsoid.from_dict(ell)
        check(lib().oc_cloth_step_physics(self.h, int(nsteps), C.byref(e) if e is not None else None,
                                          float(provot_slack)))

    # -- differentiable rollouts ---------------------------------------------
    def set_spring_stiffness(self, ks):
        k = np.ascontiguousarray(ks, np.float32)
        if k.shape != (self.S,):
            raise ValueError(f"expected {self.S} stiffness values")
        check(lib().oc_cloth_set_spring_stiffness(self.h, ptr(k)))

    def spring_stiffness(self) -> np.ndarray:
        k = np.zeros(self.S, np.float32)
        check(lib().oc_cloth_get_spring_stiffness(self.h, ptr(k)))
        return k

    def rollout_forward(self, nsteps: int, checkpoint_every: int = 16):
        """nsteps steps from the current state (dv reset), checkpointed."""
        check(lib().oc_cloth_rollout_forward(self.h, int(nsteps), int(checkpoint_every)))

    def rollout_backward(self, grad_x, grad_v=None):
        """(dL/dks, dL/dx0, dL/dv0) of the last rollout_forward given dL/dx_T."""
        gx = _vec(grad_x, self.N)
        gv = None if grad_v is None else _vec(grad_v, self.N)
        gks = np.zeros(self.S, np.float32)
        gx0 = np.zeros((self.N, 3), np.float32)
        gv0 = np.zeros((self.N, 3), np.float32)
        check(lib().oc_cloth_rollout_backward(self.h, ptr(gx), ptr(gv), ptr(gks), ptr(gx0), ptr(gv0)))
        return gks, gx0, gv0

    # -- introspection -------------------------------------------------------
    def pairs(self):
        a = np.zeros(self.P, np.int32); b = np.zeros(self.P, np.int32)
        check(lib().oc_cloth_get_pairs(self.h, ptr(a), ptr(b)))
        return a, b

    def system(self):
        """(pair blocks [P,6], rhs [N,3]) of the last assembly."""
        Cb = np.zeros((self.P, 6), np.float32); b = np.zeros((self.N, 3), np.float32)
        check(lib().oc_cloth_get_system(self.h, ptr(Cb), ptr(b)))
        return Cb, b


class VerletCloth:
    """Position-Verlet cloth (oc_verlet), the Verlet demo's step
    (OpenCloth_Verlet/OpenCloth_Verlet/main.cpp:428-562)."""

    def __init__(self, nx, ny, X, springs, pinned=None, mass=1.0, gravity=(0.0, -9.81, 0.0),
                 damping=-0.0125, dt=1.0 / 60.0, ground_clamp=True, X_last=None):
        self.nx, self.ny = int(nx), int(ny)
        self.N = self.nx * self.ny
        p = OcVerletParams()
        p.mass = mass
        p.gravity[0], p.gravity[1], p.gravity[2] = gravity
        p.damping = damping
        p.dt = dt
        p.flags = 1 if ground_clamp else 0
        self._p = p
        X = _vec(X, self.N)
        XL = None if X_last is None else _vec(X_last, self.N)
        springs = np.ascontiguousarray(springs, dtype=SPRING_DTYPE)
        pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        h = C.c_void_p()
        check(lib().oc_verlet_create(self.nx, self.ny, ptr(X), ptr(XL), ptr(springs), len(springs),
                                     ptr(pins), C.byref(p), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().oc_verlet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_state(self):
        X = np.zeros((self.N, 3), np.float32); XL = np.zeros((self.N, 3), np.float32)
        check(lib().oc_verlet_get_state(self.h, ptr(X), ptr(XL)))
        return X, XL

    def set_state(self, X=None, X_last=None):
        Xa = None if X is None else _vec(X, self.N)
        XLa = None if X_last is None else _vec(X_last, self.N)
        check(lib().oc_verlet_set_state(self.h, ptr(Xa), ptr(XLa)))

    def compute_forces(self):
        """ComputeForces(dt) (VL:448-484)."""
        F = np.zeros((self.N, 3), np.float32)
        check(lib().oc_verlet_compute_forces(self.h, ptr(F)))
        return F

    def step(self, nsteps: int = 1, ell: dict | None = None):
        """StepPhysics of the Verlet demo (VL:557-562)."""
        e = None if ell is None else OcEllipsoid.from_dict(ell)
        check(lib().oc_verlet_step(self.h, int(nsteps), C.byref(e) if e is not None else None))


class VerletGrid:
    """Grid Verlet (oc_vgrid): the GPU demos' Verlet kernel
    (OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/verlet_kernel.cu:70-197)
    on a width x height grid of float4 positions, 12-neighbour stencil."""

    def __init__(self, width, height, X, X_last=None, pinned=None, params: dict | None = None):
        from .scenes import vgrid_params

        self.width, self.height = int(width), int(height)
        self.N = self.width * self.height
        self.params = dict(params or vgrid_params(self.width, self.height))
        self._p = OcVgridParams.from_dict(self.params)
        X4 = self._vec4(X)
        XL4 = None if X_last is None else self._vec4(X_last)
        pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        if pins is not None and pins.shape != (self.N,):
            raise ValueError("pinned must have one entry per vertex")
        h = C.c_void_p()
        check(lib().oc_vgrid_create(self.width, self.height, ptr(X4), ptr(XL4), ptr(pins),
                                    C.byref(self._p), C.byref(h)))
        self.h = h

    @classmethod
    def from_scene(cls, sc) -> "VerletGrid":
        return cls(sc.width, sc.height, sc.X, sc.X_last, sc.pinned, sc.params)

    def _vec4(self, a):
        a = np.ascontiguousarray(a, np.float32)
        if a.shape != (self.N, 4):
            raise ValueError(f"expected ({self.N}, 4) positions, got {a.shape}")
        return a

    def close(self):
        if getattr(self, "h", None):
            lib().oc_vgrid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: dict):
        p = OcVgridParams.from_dict(params)
        check(lib().oc_vgrid_set_params(self.h, C.byref(p)))
        self.params, self._p = dict(params), p

    def set_stream(self, stream):
        check(lib().oc_vgrid_set_stream(self.h, stream))

    def get_state(self):
        X = np.zeros((self.N, 4), np.float32)
        XL = np.zeros((self.N, 4), np.float32)
        check(lib().oc_vgrid_get_state(self.h, ptr(X), ptr(XL)))
        return X, XL

    def set_state(self, X=None, X_last=None):
        Xa = None if X is None else self._vec4(X)
        XLa = None if X_last is None else self._vec4(X_last)
        check(lib().oc_vgrid_set_state(self.h, ptr(Xa), ptr(XLa)))

    def step(self, nsteps: int = 1, vbo: int | None = None):
        """nsteps kernel steps; vbo: device address of N float4 that receives
        the last step's positions (the demo's mapped vertex buffer)."""
        check(lib().oc_vgrid_step(self.h, int(nsteps), vbo))

    def synchronize(self):
        check(lib().oc_vgrid_synchronize(self.h))


class PBDCloth:
    """Position-based-dynamics cloth (oc_pbd), the PBD demo's step
    (OpenCloth_PositionBasedDynamics/.../main.cpp:601-925) over given
    distance and bending constraint lists."""

    def __init__(self, X, distance, bending, pinned=None, params: PBDParams | None = None, V=None):
        X = np.ascontiguousarray(X, np.float32)
        self.N = len(X)
        X = _vec(X, self.N)
        self.params = params or PBDParams()
        p = OcPbdParams()
        p.mass = self.params.mass
        p.gravity[0], p.gravity[1], p.gravity[2] = self.params.gravity
        p.dt = self.params.dt
        p.global_damping = self.params.global_damping
        p.k_damp = self.params.k_damp
        p.solver_iterations = self.params.solver_iterations
        p.flags = ((_lib.OC_FLAG_GROUND_CLAMP if self.params.ground else 0)
                   | (_lib.OC_FLAG_LIST_ORDER if getattr(self.params, "list_order", False) else 0))
        self._p = p
        dist = np.ascontiguousarray(distance, dtype=PBD_DISTANCE_DTYPE)
        bend = np.ascontiguousarray(bending, dtype=PBD_BENDING_DTYPE)
        pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        if pins is not None and pins.shape != (self.N,):
            raise ValueError("pinned must have one entry per particle")
        h = C.c_void_p()
        check(lib().oc_pbd_create(self.N, ptr(X), ptr(pins), ptr(dist), len(dist), ptr(bend),
                                  len(bend), C.byref(p), C.byref(h)))
        self.h = h
        if V is not None:
            self.set_state(V=V)

    @classmethod
    def from_scene(cls, sc) -> "PBDCloth":
        return cls(sc.X, sc.distance, sc.bending, sc.pinned, sc.params, sc.V)

    def close(self):
        if getattr(self, "h", None):
            lib().oc_pbd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, X=None, V=None):
        Xa = None if X is None else _vec(X, self.N)
        Va = None if V is None else _vec(V, self.N)
        check(lib().oc_pbd_set_state(self.h, ptr(Xa), ptr(Va)))

    def get_state(self):
        X = np.zeros((self.N, 3), np.float32); V = np.zeros((self.N, 3), np.float32)
        check(lib().oc_pbd_get_state(self.h, ptr(X), ptr(V)))
        return X, V

    def num_colors(self):
        a = C.c_int32(); b = C.c_int32()
        check(lib().oc_pbd_num_colors(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def step(self, nsteps: int = 1, ell: dict | None = None):
        """StepPhysics of the PBD demo (PBD:915-925)."""
        e = None if ell is None else OcEllipsoid.from_dict(ell)
        check(lib().oc_pbd_step(self.h, int(nsteps), C.byref(e) if e is not None else None))

    def synchronize(self):
        check(lib().oc_pbd_synchronize(self.h))
