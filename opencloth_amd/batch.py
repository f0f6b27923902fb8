"""Batched cloths over the C ABI (oc_batch_*): many independent cloths of
uneven sizes stepped by one persistent kernel launch per step, replayed from a
captured HIP graph."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import OcBatchSchedule, check, lib, ptr
from .cloth import to_oc_params
from .scenes import SPRING_DTYPE, BatchScene, StepParams


class ClothBatch:
    def __init__(self, nx, ny, X, springs, n_springs, pinned=None, params: StepParams | None = None,
                 V=None):
        self.nx = np.ascontiguousarray(nx, np.int32)
        self.ny = np.ascontiguousarray(ny, np.int32)
        self.n = len(self.nx)
        self.N = int((self.nx.astype(np.int64) * self.ny).sum())
        self.params = params or StepParams()
        X = np.ascontiguousarray(X, np.float32).reshape(self.N, 3)
        V = None if V is None else np.ascontiguousarray(V, np.float32).reshape(self.N, 3)
        springs = np.ascontiguousarray(springs, SPRING_DTYPE)
        ns = np.ascontiguousarray(n_springs, np.int32)
        pins = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        self._p = to_oc_params(self.params)
        h = C.c_void_p()
        check(lib().oc_batch_create(self.n, ptr(self.nx), ptr(self.ny), ptr(X), ptr(V), ptr(springs),
                                    ptr(ns), ptr(pins), C.byref(self._p), C.byref(h)))
        self.h = h
        self.n_springs_total = int(ns.astype(np.int64).sum())

    @classmethod
    def from_scene(cls, sc: BatchScene) -> "ClothBatch":
        return cls(sc.nx, sc.ny, sc.X, sc.springs, sc.n_springs, sc.pinned, sc.params, sc.V)

    def close(self):
        if getattr(self, "h", None):
            lib().oc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: StepParams):
        self.params = params
        self._p = to_oc_params(params)
        check(lib().oc_batch_set_params(self.h, C.byref(self._p)))

    def set_stream(self, stream_handle: int | None):
        check(lib().oc_batch_set_stream(self.h, C.c_void_p(stream_handle or 0)))

    def schedule(self):
        s = OcBatchSchedule()
        check(lib().oc_batch_get_schedule(self.h, C.byref(s)))
        return s.block_threads, s.vertices_per_thread

    def set_schedule(self, block_threads: int, vertices_per_thread: int):
        s = OcBatchSchedule(block_threads, vertices_per_thread)
        check(lib().oc_batch_set_schedule(self.h, C.byref(s)))

    def autotune(self):
        s = OcBatchSchedule()
        check(lib().oc_batch_autotune(self.h, C.byref(s)))
        return s.block_threads, s.vertices_per_thread

    def threads_per_cloth(self) -> np.ndarray:
        t = np.zeros(self.n, np.int32)
        check(lib().oc_batch_threads_per_cloth(self.h, ptr(t)))
        return t

    def set_state(self, X=None, V=None):
        Xa = None if X is None else np.ascontiguousarray(X, np.float32)
        Va = None if V is None else np.ascontiguousarray(V, np.float32)
        check(lib().oc_batch_set_state(self.h, ptr(Xa), ptr(Va)))

    def get_state(self):
        X = np.zeros((self.N, 3), np.float32)
        V = np.zeros((self.N, 3), np.float32)
        check(lib().oc_batch_get_state(self.h, ptr(X), ptr(V)))
        return X, V

    def step(self, nsteps: int = 1, graph: bool = True):
        f = lib().oc_batch_step if graph else lib().oc_batch_step_eager
        check(f(self.h, int(nsteps)))

    def status(self):
        it = np.zeros(self.n, np.int32)
        st = np.zeros(self.n, np.int32)
        check(lib().oc_batch_get_status(self.h, ptr(it), ptr(st)))
        return it, st

    def sticky_status(self):
        """Per cloth, the OR of the status bits of every step since the last
        call (then cleared)."""
        st = np.zeros(self.n, np.int32)
        check(lib().oc_batch_sticky_status(self.h, ptr(st)))
        return st

    def synchronize(self):
        check(lib().oc_batch_synchronize(self.h))

    # -- differentiable rollouts ---------------------------------------------
    def set_spring_stiffness(self, ks):
        k = np.ascontiguousarray(ks, np.float32)
        check(lib().oc_batch_set_spring_stiffness(self.h, ptr(k)))

    def spring_stiffness(self, n_springs: int) -> np.ndarray:
        k = np.zeros(n_springs, np.float32)
        check(lib().oc_batch_get_spring_stiffness(self.h, ptr(k)))
        return k

    def rollout_forward(self, nsteps: int, checkpoint_every: int = 10):
        check(lib().oc_batch_rollout_forward(self.h, int(nsteps), int(checkpoint_every)))

    def rollout_backward(self, grad_x, grad_v=None, n_springs: int | None = None):
        """(dL/dks, dL/dx0, dL/dv0) of the last rollout_forward given dL/dx_T."""
        gx = np.ascontiguousarray(grad_x, np.float32).reshape(self.N, 3)
        gv = None if grad_v is None else np.ascontiguousarray(grad_v, np.float32).reshape(self.N, 3)
        gks = np.zeros(n_springs if n_springs is not None else self.n_springs_total, np.float32)
        gx0 = np.zeros((self.N, 3), np.float32)
        gv0 = np.zeros((self.N, 3), np.float32)
        check(lib().oc_batch_rollout_backward(self.h, ptr(gx), ptr(gv), ptr(gks), ptr(gx0), ptr(gv0)))
        return gks, gx0, gv0
