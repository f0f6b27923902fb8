"""Shared helpers of the test-suite: oracle parameters that mirror a GPU
configuration, and error metrics."""
from __future__ import annotations

import sys

import numpy as np

from opencloth_amd.scenes import StepParams

def _tile_shape():
    """The single-cloth kernels' tile shape, as the library reports it (the
    reduction order of the CG dots depends on it); host only, no GPU call."""
    from opencloth_amd._lib import tile_shape

    return tile_shape()


TILE_X, TILE_Y = _tile_shape()


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
