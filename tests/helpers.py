"""Shared helpers of the test-suite: oracle parameters that mirror a GPU
configuration, and error metrics."""
from __future__ import annotations

import sys

import numpy as np

from opencloth_amd.scenes import StepParams

# tile shape of the HIP implicit kernels (opencloth_amd/csrc/oc_internal.hpp)
TILE_X, TILE_Y = 32, 8


def oracle_params(p: StepParams, nx: int, ny: int, gpu_order: bool = True):
    """oracle Params for StepParams p; gpu_order selects the HIP reduction tree."""
    from oracle import make_params

    if gpu_order:
        return make_params(p, nx=nx, ny=ny, tx=TILE_X, ty=TILE_Y, dot_mode=1)
    return make_params(p, nx=nx, ny=ny, dot_mode=0)


def rel_pos_err(X, Xref) -> float:
    """max_i |x_i - x_i^ref| / max_i |x_i^ref| (norm-wise; components cross 0)."""
    X = np.asarray(X, np.float64); Xref = np.asarray(Xref, np.float64)
    return float(np.linalg.norm(X - Xref, axis=-1).max() / np.linalg.norm(Xref, axis=-1).max())


def say(*a):
    """Print and flush, so a GPU run that dies still names its case."""
    print(*a, flush=True)
    sys.stdout.flush()
