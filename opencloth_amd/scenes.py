"""Cloth scenes: grid positions, the reference spring list, pins and the
workload configurations, as numpy arrays ready for the C ABI.

Everything here is host-side input generation.  The arithmetic of the grid
and the rest lengths is float32 in the reference's expression order so the
arrays equal what the reference's InitGL builds:

* positions  -- OpenCloth_ImplicitEuler/.../main.cpp:462-466 (IE)
* springs    -- IE:496-530 with AddSpring IE:336-346
* constants  -- IE:309-316 (the reference stores the damping Kd negative)

Random inputs use SplitMix64 with Box-Muller, so any host can regenerate the
same scene from a seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Spring record in the reference layout (IE:68-73): 24 bytes.
SPRING_DTYPE = np.dtype(
    [("p1", "<i4"), ("p2", "<i4"), ("rest_length", "<f4"), ("Ks", "<f4"), ("Kd", "<f4"), ("type", "<i4")]
)
STRUCTURAL, SHEAR, BEND = 0, 1, 2


# ---------------------------------------------------------------------------
# random numbers
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from SplitMix64 (top 53 bits)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller from uniform(seed, 2*ceil(n/2))."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[:m]  # (0, 1]
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


# ---------------------------------------------------------------------------
# grid and springs
# ---------------------------------------------------------------------------
def grid_positions(nx: int, ny: int, fullsize: float = 4.0, y0: float = 0.0) -> np.ndarray:
    """(N,3) float32; vertex k = j*nx + i at
    (((i/(nx-1))*2-1)*halfsize, y0, (j/(ny-1))*fullsize)   (IE:462-466,
    which uses y0 = fullsize+1)."""
    f32 = np.float32
    fs = f32(fullsize)
    half = fs / f32(2.0)
    i = np.arange(nx, dtype=np.float32)
    j = np.arange(ny, dtype=np.float32)
    xi = ((i / f32(nx - 1)) * f32(2) - f32(1)) * half
    zj = (j / f32(ny - 1)) * fs
    X = np.empty((ny, nx, 3), np.float32)
    X[:, :, 0] = xi[None, :]
    X[:, :, 1] = f32(y0)
    X[:, :, 2] = zj[:, None]
    return X.reshape(-1, 3)


def grid_pairs(nx: int, ny: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(p1, p2, type) of the reference spring list (IE:496-530), in order:
    horizontal, vertical, two shears per quad, bend rows (each row's last bend
    twice), bend columns (each column's last bend twice)."""
    u, v = nx, ny
    if u < 3 or v < 3:
        raise ValueError("the reference spring topology needs at least 3x3 vertices")
    ps, qs, ts = [], [], []
    # horizontal: for l1 < v, l2 < u-1
    l1, l2 = np.meshgrid(np.arange(v), np.arange(u - 1), indexing="ij")
    ps.append((l1 * u + l2).ravel()); qs.append((l1 * u + l2 + 1).ravel()); ts.append(STRUCTURAL)
    # vertical: for l1 < u, l2 < v-1
    l1, l2 = np.meshgrid(np.arange(u), np.arange(v - 1), indexing="ij")
    ps.append((l2 * u + l1).ravel()); qs.append(((l2 + 1) * u + l1).ravel()); ts.append(STRUCTURAL)
    # shear: two per quad, interleaved
    l1, l2 = np.meshgrid(np.arange(v - 1), np.arange(u - 1), indexing="ij")
    a1 = (l1 * u + l2).ravel(); b1 = ((l1 + 1) * u + l2 + 1).ravel()
    a2 = ((l1 + 1) * u + l2).ravel(); b2 = (l1 * u + l2 + 1).ravel()
    ps.append(np.stack([a1, a2], 1).ravel()); qs.append(np.stack([b1, b2], 1).ravel()); ts.append(SHEAR)
    # bend rows: per row u-2 springs then the duplicate of the last
    l1, l2 = np.meshgrid(np.arange(v), np.arange(u - 2), indexing="ij")
    a = l1 * u + l2; b = l1 * u + l2 + 2
    a = np.concatenate([a, (np.arange(v) * u + (u - 3))[:, None]], 1)
    b = np.concatenate([b, (np.arange(v) * u + (u - 1))[:, None]], 1)
    ps.append(a.ravel()); qs.append(b.ravel()); ts.append(BEND)
    # bend columns
    l1, l2 = np.meshgrid(np.arange(u), np.arange(v - 2), indexing="ij")
    a = l2 * u + l1; b = (l2 + 2) * u + l1
    a = np.concatenate([a, ((v - 3) * u + np.arange(u))[:, None]], 1)
    b = np.concatenate([b, ((v - 1) * u + np.arange(u))[:, None]], 1)
    ps.append(a.ravel()); qs.append(b.ravel()); ts.append(BEND)
    p1 = np.concatenate(ps).astype(np.int32)
    p2 = np.concatenate(qs).astype(np.int32)
    ty = np.concatenate([np.full(len(x), t, np.int32) for x, t in zip(ps, ts)])
    return p1, p2, ty


def rest_lengths(X: np.ndarray, p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
    """sqrt(dot(dp,dp)) in float32, dp = X[p1]-X[p2] (IE:343-344)."""
    d = X[p1].astype(np.float32) - X[p2].astype(np.float32)
    s = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    s = s + d[:, 2] * d[:, 2]
    return np.sqrt(s.astype(np.float32))


def grid_springs(nx, ny, X, ks=(50.0, 50.0, 50.0), kd=(-0.25, -0.25, -0.25)) -> np.ndarray:
    """Reference spring list as a SPRING_DTYPE array.  ks/kd per type
    (structural, shear, bend); kd uses the reference sign (negative damps)."""
    p1, p2, ty = grid_pairs(nx, ny)
    s = np.zeros(len(p1), SPRING_DTYPE)
    s["p1"], s["p2"], s["type"] = p1, p2, ty
    s["rest_length"] = rest_lengths(X, p1, p2)
    s["Ks"] = np.asarray(ks, np.float32)[ty]
    s["Kd"] = np.asarray(kd, np.float32)[ty]
    return s


# ---------------------------------------------------------------------------
# step parameters and configurations
# ---------------------------------------------------------------------------
TOL_ABSOLUTE, TOL_RELATIVE, TOL_NONE = 0, 1, 2


@dataclass
class StepParams:
    mass: float = 1.0
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1.0 / 60.0
    cg_max_iters: int = 10
    cg_tol: float = 1e-3
    cg_tol_mode: int = TOL_ABSOLUTE
    warm_start: bool = True
    ground_clamp: bool = False
    cg_persistent: bool = False  # all CG iterations in one cooperative launch
    cg_jacobi: bool = False      # Jacobi-preconditioned CG (Baraff-Witkin modified PCG)
    damping: float = 0.0         # velocity damping c: force c*v on free vertices (BW:708)
    cg_single_reduction: bool = False  # one reduction point per CG iteration (Chronopoulos-Gear)


@dataclass
class Scene:
    name: str
    nx: int
    ny: int
    X: np.ndarray  # (N,3) float32
    V: np.ndarray  # (N,3) float32
    springs: np.ndarray  # SPRING_DTYPE
    pinned: np.ndarray  # (N,) uint8
    params: StepParams
    steps: int
    extra: dict = field(default_factory=dict)

    @property
    def n_vertices(self) -> int:
        return self.nx * self.ny


def perturb(X: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    """X + N(0, sigma) per coordinate, float32."""
    z = normal(seed, X.size).reshape(X.shape)
    return (X.astype(np.float64) + sigma * z).astype(np.float32)


def plain_scene(n: int = 21, steps: int = 300) -> Scene:
    """21x21 over a 4x4 square at y=0, ks 50, kd 0.25, mass 1, g -9.81,
    dt 1/60, two top corners pinned, CG 10 iterations with absolute tol 1e-3,
    300 steps from rest."""
    X = grid_positions(n, n, 4.0, 0.0)
    springs = grid_springs(n, n, X)
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    p = StepParams(dt=1.0 / 60.0, cg_max_iters=10, cg_tol=1e-3, cg_tol_mode=TOL_ABSOLUTE)
    return Scene(f"plain_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, steps)


def large_scene(n: int = 512, steps: int = 100, seed: int = 0) -> Scene:
    """512x512 over the 4x4 square, positions perturbed by N(0, 1e-3*spacing)
    from `seed` (rest lengths from the unperturbed grid), top edge (row 0)
    pinned, dt 1/120, CG fixed at 30 iterations."""
    X0 = grid_positions(n, n, 4.0, 0.0)
    springs = grid_springs(n, n, X0)
    spacing = 4.0 / (n - 1)
    X = perturb(X0, 1e-3 * spacing, seed)
    pinned = np.zeros(n * n, np.uint8)
    pinned[:n] = 1
    p = StepParams(dt=1.0 / 120.0, cg_max_iters=30, cg_tol=0.0, cg_tol_mode=TOL_NONE)
    return Scene(f"large_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, steps)


@dataclass
class BatchScene:
    name: str
    nx: np.ndarray  # (C,) int32
    ny: np.ndarray
    X: np.ndarray  # (sum N, 3) float32, cloths concatenated
    V: np.ndarray
    springs: np.ndarray  # concatenated SPRING_DTYPE, per-cloth local indices
    n_springs: np.ndarray  # (C,) int32
    pinned: np.ndarray  # (sum N,) uint8
    params: StepParams
    steps: int
    extra: dict = field(default_factory=dict)

    @property
    def n_cloths(self) -> int:
        return len(self.nx)

    @property
    def voff(self) -> np.ndarray:
        n = (self.nx.astype(np.int64) * self.ny)
        return np.concatenate([[0], np.cumsum(n)[:-1]])

    @property
    def soff(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n_springs.astype(np.int64))[:-1]])

    def cloth(self, c: int) -> Scene:
        """Cloth c as a single-cloth Scene."""
        v0, s0 = int(self.voff[c]), int(self.soff[c])
        n = int(self.nx[c] * self.ny[c])
        return Scene(f"{self.name}[{c}]", int(self.nx[c]), int(self.ny[c]), self.X[v0:v0 + n],
                     self.V[v0:v0 + n], self.springs[s0:s0 + int(self.n_springs[c])],
                     self.pinned[v0:v0 + n], self.params, self.steps)


def batch_scene(n_cloths: int = 4096, sides=(16, 128), seed: int = 0, steps: int = 200,
                ks_range=(20.0, 200.0), kd_range=(0.1, 0.5), pins=(2, 8),
                dt: float = 1.0 / 60.0, cg_max_iters: int = 40, cg_tol: float = 1e-4) -> BatchScene:
    """n_cloths independent square cloths over the 4x4 square at y=0; side
    uniform in [sides], ks log-uniform in ks_range and kd uniform in kd_range
    per cloth, `pins` random top-edge (row 0) vertices pinned per cloth; CG at
    most cg_max_iters iterations with per-cloth relative tolerance cg_tol."""
    u = uniform(seed, 4 * n_cloths).reshape(4, n_cloths)
    side = (sides[0] + np.floor(u[0] * (sides[1] - sides[0] + 1))).astype(np.int32)
    ks = np.exp(np.log(ks_range[0]) + u[1] * (np.log(ks_range[1]) - np.log(ks_range[0])))
    kd = kd_range[0] + u[2] * (kd_range[1] - kd_range[0])
    npin = (pins[0] + np.floor(u[3] * (pins[1] - pins[0] + 1))).astype(np.int32)
    cache = {}  # side -> (positions, springs): the topology depends on the side only
    for sd in np.unique(side):
        X0 = grid_positions(int(sd), int(sd), 4.0, 0.0)
        cache[int(sd)] = (X0, grid_springs(int(sd), int(sd), X0))
    nverts = side.astype(np.int64) ** 2
    nspr = np.array([len(cache[int(sd)][1]) for sd in side], np.int64)
    voff = np.concatenate([[0], np.cumsum(nverts)])
    soff = np.concatenate([[0], np.cumsum(nspr)])
    X = np.empty((int(voff[-1]), 3), np.float32)
    springs = np.empty(int(soff[-1]), SPRING_DTYPE)
    pinned = np.zeros(int(voff[-1]), np.uint8)
    for c in range(n_cloths):
        X0, sp0 = cache[int(side[c])]
        X[voff[c]:voff[c + 1]] = X0
        springs[soff[c]:soff[c + 1]] = sp0
        perm = np.argsort(uniform(seed * 1000003 + c + 1, int(side[c])))
        pinned[voff[c] + perm[: int(npin[c])]] = 1
    springs["Ks"] = np.repeat(ks.astype(np.float32), nspr)
    springs["Kd"] = np.repeat(-kd.astype(np.float32), nspr)
    p = StepParams(dt=dt, cg_max_iters=cg_max_iters, cg_tol=cg_tol, cg_tol_mode=TOL_RELATIVE)
    return BatchScene(f"batch_{n_cloths}", side.copy(), side.copy(), X, np.zeros_like(X),
                      springs, nspr.astype(np.int32), pinned, p, steps)


def smooth_target(X: np.ndarray, seed: int, amplitude: float = 0.2, n_waves: int = 4) -> np.ndarray:
    """X displaced by a sum of n_waves random sinusoids: wave k has a unit
    displacement direction d_k, wave vector q_k in the cloth plane (|q_k| in
    [0.5, 2]) and phase p_k; each has the given amplitude."""
    u = uniform(seed, 6 * n_waves).reshape(n_waves, 6)
    out = X.astype(np.float64).copy()
    for k in range(n_waves):
        th, ph = 2 * math.pi * u[k, 0], math.acos(2 * u[k, 1] - 1)
        d = np.array([math.sin(ph) * math.cos(th), math.cos(ph), math.sin(ph) * math.sin(th)])
        ang, mag = 2 * math.pi * u[k, 2], 0.5 + 1.5 * u[k, 3]
        q = mag * np.array([math.cos(ang), 0.0, math.sin(ang)])
        phase = 2 * math.pi * u[k, 4]
        out += amplitude * np.sin(X.astype(np.float64) @ q + phase)[:, None] * d[None, :]
    return out.astype(np.float32)


def diff_scene(n: int = 128, steps: int = 120, seed: int = 0, ks: float = 50.0) -> Scene:
    """Differentiable rollout: n x n cloth (top edge pinned), dt 1/60, CG fixed
    at 40 iterations, per-spring ks (all `ks` to start); target = rest grid
    displaced by smooth_target(seed).  Target in extra['target']."""
    X = grid_positions(n, n, 4.0, 0.0)
    springs = grid_springs(n, n, X, ks=(ks, ks, ks))
    pinned = np.zeros(n * n, np.uint8)
    pinned[:n] = 1
    p = StepParams(dt=1.0 / 60.0, cg_max_iters=40, cg_tol=0.0, cg_tol_mode=TOL_NONE)
    return Scene(f"diff_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, steps,
                 extra={"target": smooth_target(X, seed)})


def batch_diff_scene(n_cloths: int = 512, sides=(32, 96), steps: int = 60, seed: int = 0,
                     ks_range=(20.0, 200.0), cg_tol: float = 1e-6) -> BatchScene:
    """Batched differentiable rollouts: n_cloths cloths (top edge pinned),
    sides uniform in [sides], per-spring ks log-uniform in ks_range, dt 1/60,
    CG at most 40 iterations with relative tolerance cg_tol; per-cloth smooth
    targets (smooth_target, seed + 1 + cloth).  Targets in `extra`."""
    b = batch_scene(n_cloths=n_cloths, sides=sides, seed=seed, steps=steps, cg_tol=cg_tol)
    u = uniform(seed + 99991, len(b.springs))
    b.springs["Ks"] = np.exp(np.log(ks_range[0]) + u * (np.log(ks_range[1]) - np.log(ks_range[0]))).astype(np.float32)
    b.pinned[:] = 0
    targets = np.empty_like(b.X)
    for c in range(n_cloths):
        v0, n = int(b.voff[c]), int(b.nx[c])
        b.pinned[v0:v0 + n] = 1
        targets[v0:v0 + n * n] = smooth_target(b.X[v0:v0 + n * n], seed + 1 + c)
    b.name = f"batch_diff_{n_cloths}"
    b.extra = {"target": targets}
    return b


def ellipsoid(translate=(0.0, 2.0, 0.0), angle_deg: float = 45.0, radius: float = 1.0) -> dict:
    """The reference's collider (IE:541-544): translate * rotate(angle about x)
    * scale(r, r, r/2), as column-major float32 4x4 matrices with the inverse;
    object-space center 0 and radius 1 (IE:85-86).  Built here in float64 and
    rounded, not with GLM's float matrix code."""
    a = math.radians(angle_deg)
    T = np.eye(4); T[:3, 3] = translate
    R = np.eye(4); R[1, 1] = R[2, 2] = math.cos(a); R[1, 2] = -math.sin(a); R[2, 1] = math.sin(a)
    S = np.diag([radius, radius, radius / 2, 1.0])
    M = T @ R @ S
    return {"M": M.T.astype(np.float32).ravel(), "Minv": np.linalg.inv(M).T.astype(np.float32).ravel(),
            "center": np.zeros(3, np.float32), "radius": 1.0}


def demo_scene(n: int = 21) -> Scene:
    """The implicit demo's own scene (IE:59-62, IE:309-318, IE:434-546): 21x21
    at height fullsize+1 = 5, Ks 0.75/0.75/0.95, Kd -0.25, mass 0.5, gravity
    force -0.00981 (here acceleration -0.01962, so mass*g is the same float),
    dt 1/60, CG 10 iterations with absolute tol 1e-3, ground clamp, pins
    {0, numX}; the post-step passes use ellipsoid() and Provot slack 1."""
    X = grid_positions(n, n, 4.0, 5.0)
    springs = grid_springs(n, n, X, ks=(0.75, 0.75, 0.95), kd=(-0.25, -0.25, -0.25))
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    g = float(np.float32(-0.00981) * np.float32(2.0))
    p = StepParams(mass=0.5, gravity=(0.0, g, 0.0), dt=1.0 / 60.0, cg_max_iters=10, cg_tol=1e-3,
                   cg_tol_mode=TOL_ABSOLUTE, ground_clamp=True)
    return Scene(f"demo_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, 600,
                 extra={"ellipsoid": ellipsoid(), "provot_slack": 1.0})


def bw_demo_scene(n: int = 21) -> Scene:
    """The Baraff-Witkin demo's scene (OpenCloth_Implicit/.../main.cpp,
    BW:60-128, BW:495-590): 21x21 at height fullsize+1 = 5 (BW:500), Ks
    0.75/0.75/0.95, Kd -0.25 (BW:105-107), mass 1, gravity force -0.00981
    (BW:108-109; an acceleration at mass 1), dt 5/60 (BW:65), velocity
    damping DEFAULT_DAMPING -0.125 on every free vertex (BW:706-708), 10 CG
    iterations to the relative tolerance 1e-3 (BW:126-128, BW:319) with the
    Jacobi preconditioner BW:784-785 means, ground clamp (BW:794-796), pins
    {0, numX}; the post-step passes use the same collider (BW:583-586) and
    Provot slack 1.01 (BW:808).  BW's own step multiplies dv by dt twice
    (BW:791) and leaves the damping out of its Jacobian; neither is copied."""
    X = grid_positions(n, n, 4.0, 5.0)
    springs = grid_springs(n, n, X, ks=(0.75, 0.75, 0.95), kd=(-0.25, -0.25, -0.25))
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    p = StepParams(mass=1.0, gravity=(0.0, -0.00981, 0.0), dt=5.0 / 60.0, cg_max_iters=10, cg_tol=1e-3,
                   cg_tol_mode=TOL_RELATIVE, ground_clamp=True, cg_jacobi=True, damping=-0.125)
    return Scene(f"bw_demo_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, 600,
                 extra={"ellipsoid": ellipsoid(), "provot_slack": 1.01})


def verlet_demo_scene(n: int = 21) -> Scene:
    """The Verlet demo's scene (VL:59-106, VL:240-328): height 5, Ks 50.75 /
    50.75 / 50.95, Kd -0.25, mass 1, gravity -0.00981, damping -0.0125,
    dt 1/60, ground clamp, pins {0, numX}, ellipsoid()."""
    X = grid_positions(n, n, 4.0, 5.0)
    springs = grid_springs(n, n, X, ks=(50.75, 50.75, 50.95), kd=(-0.25, -0.25, -0.25))
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    p = StepParams(mass=1.0, gravity=(0.0, -0.00981, 0.0), dt=1.0 / 60.0, ground_clamp=True)
    return Scene(f"verlet_demo_{n}x{n}", n, n, X, np.zeros_like(X), springs, pinned, p, 600,
                 extra={"ellipsoid": ellipsoid(), "damping": -0.0125})


# ---------------------------------------------------------------------------
# grid Verlet: the GPU demos' kernel
# (OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/, CUK =
# verlet_kernel.cu, CUM = main.cpp)
# ---------------------------------------------------------------------------
@dataclass
class VgridScene:
    name: str
    width: int
    height: int
    X: np.ndarray        # (N, 4) float32, w = 1
    X_last: np.ndarray   # (N, 4) float32
    pinned: np.ndarray   # (N,) uint8
    params: dict         # mass, gravity, damping, dt, spacing, ks, kd
    steps: int


def vgrid_params(n_x: int, n_y: int, size: float = 4.0) -> dict:
    """The demo's constants: mass 1 (CUM:100), gravity (0, -0.00981, 0)
    (CUK:123), DEFAULT_DAMPING -0.0125 (CUM:95), dt 1/60 (CUM:102), spacing
    (sizeX/numX, sizeY/numY) in FP32 (CUM:601), Ks 50.75 / 50.75 / 50.95 and
    Kd -0.25 (CUK:8-10)."""
    f32 = np.float32
    return {"mass": 1.0, "gravity": (0.0, float(f32(-0.00981)), 0.0), "damping": float(f32(-0.0125)),
            "dt": float(f32(1.0) / f32(60.0)),
            "spacing": (float(f32(size) / f32(n_x - 1)), float(f32(size) / f32(n_y - 1))),
            "ks": (float(f32(50.75)), float(f32(50.75)), float(f32(50.95))),
            "kd": (-0.25, -0.25, -0.25)}


def vgrid_demo_scene(n: int = 64) -> VgridScene:
    """The GPU demo's cloth (CUM:56, CUM:436-444): n x n texels (numX = n-1),
    vertex j*n + i at (((i/(n-1))*2-1)*2, 5, (j/(n-1))*4, 1), at rest, pins
    0 and n-1 (CUK:126)."""
    X3 = grid_positions(n, n, 4.0, 5.0)
    X = np.concatenate([X3, np.ones((n * n, 1), np.float32)], axis=1)
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    return VgridScene(f"vgrid_demo_{n}x{n}", n, n, X, X.copy(), pinned, vgrid_params(n, n), 600)


def vgrid_large_scene(n: int = 512, seed: int = 0) -> VgridScene:
    """The demo's cloth and constants at n x n, positions perturbed by
    N(0, 1e-3*spacing) (x_last = the perturbed positions)."""
    sc = vgrid_demo_scene(n)
    spacing = 4.0 / (n - 1)
    X = sc.X.copy()
    X[:, :3] = (X[:, :3] + (1e-3 * spacing) * normal(seed, 3 * n * n).reshape(-1, 3)).astype(np.float32)
    return VgridScene(f"vgrid_{n}x{n}", n, n, X, X.copy(), sc.pinned, sc.params, 100)


# ---------------------------------------------------------------------------
# position-based dynamics (OpenCloth_PositionBasedDynamics/.../main.cpp, PBD)
# ---------------------------------------------------------------------------
# AddDistanceConstraint(a, b, k) (PBD:148-162) and the triangle form of
# AddBendingConstraint(a, b, c, k) (PBD:164-178) as records.
PBD_DISTANCE_DTYPE = np.dtype([("p1", "<i4"), ("p2", "<i4"), ("k", "<f4")])
PBD_BENDING_DTYPE = np.dtype([("p1", "<i4"), ("p2", "<i4"), ("p3", "<i4"), ("k", "<f4")])


@dataclass
class PBDParams:
    mass: float = 1.0 / 441.0       # per particle (PBD:114); inverse mass 1/mass, 0 if pinned
    gravity: tuple = (0.0, -0.00981, 0.0)  # PBD:112
    dt: float = 1.0 / 60.0          # PBD:74
    global_damping: float = 0.98    # PBD:78
    k_damp: float = 0.00125         # PBD:111
    solver_iterations: int = 2      # PBD:107
    ground: bool = True             # GroundCollision in the solver loop (PBD:887)
    list_order: bool = False        # the reference's sequential order (PBD:876-889), one workgroup


@dataclass
class PBDScene:
    name: str
    nx: int
    ny: int
    X: np.ndarray
    V: np.ndarray
    distance: np.ndarray  # PBD_DISTANCE_DTYPE
    bending: np.ndarray   # PBD_BENDING_DTYPE
    pinned: np.ndarray
    params: PBDParams
    steps: int
    extra: dict = field(default_factory=dict)

    @property
    def n_vertices(self) -> int:
        return self.nx * self.ny


def pbd_grid_constraints(nx: int, ny: int, k_stretch: float = 0.25, k_bend: float = 0.5):
    """The reference's constraint lists in its order (PBD:399-433): horizontal,
    vertical and the two shear distance constraints per quad, then the
    triangle bending constraints along columns and along rows
    (getIndex(i, j) = j*nx + i, PBD:306-308)."""
    u, v = nx, ny
    l1, l2 = np.meshgrid(np.arange(v), np.arange(u - 1), indexing="ij")
    hor = np.stack([l1 * u + l2, l1 * u + l2 + 1], -1).reshape(-1, 2)
    l1, l2 = np.meshgrid(np.arange(u), np.arange(v - 1), indexing="ij")
    ver = np.stack([l2 * u + l1, (l2 + 1) * u + l1], -1).reshape(-1, 2)
    l1, l2 = np.meshgrid(np.arange(v - 1), np.arange(u - 1), indexing="ij")
    sh = np.stack([l1 * u + l2, (l1 + 1) * u + l2 + 1, (l1 + 1) * u + l2, l1 * u + l2 + 1], -1).reshape(-1, 2)
    pairs = np.concatenate([hor, ver, sh])
    dist = np.zeros(len(pairs), PBD_DISTANCE_DTYPE)
    dist["p1"], dist["p2"], dist["k"] = pairs[:, 0], pairs[:, 1], k_stretch
    i, j = np.meshgrid(np.arange(u), np.arange(v - 2), indexing="ij")
    bv = np.stack([j * u + i, (j + 1) * u + i, (j + 2) * u + i], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(u - 2), np.arange(v), indexing="ij")
    bh = np.stack([j * u + i, j * u + i + 1, j * u + i + 2], -1).reshape(-1, 3)
    tri = np.concatenate([bv, bh])
    bend = np.zeros(len(tri), PBD_BENDING_DTYPE)
    bend["p1"], bend["p2"], bend["p3"], bend["k"] = tri[:, 0], tri[:, 1], tri[:, 2], k_bend
    return dist, bend


def pbd_demo_scene(n: int = 21) -> PBDScene:
    """The PBD demo (PBD:66-134, PBD:310-464): n x n grid at height 5 over the
    4x4 square, kStretch 0.25, kBend 0.5, kDamp 0.00125, global damping 0.98,
    2 solver iterations, mass 1/N, gravity -0.00981, dt 1/60, pins {0, numX},
    the ellipsoid collider."""
    X = grid_positions(n, n, 4.0, 5.0)
    dist, bend = pbd_grid_constraints(n, n, 0.25, 0.5)
    pinned = np.zeros(n * n, np.uint8)
    pinned[0] = pinned[n - 1] = 1
    p = PBDParams(mass=float(np.float32(1.0) / np.float32(n * n)))
    return PBDScene(f"pbd_demo_{n}x{n}", n, n, X, np.zeros_like(X), dist, bend, pinned, p, 600,
                    extra={"ellipsoid": ellipsoid()})


def pbd_large_scene(n: int = 512, seed: int = 0) -> PBDScene:
    """The PBD demo's constraints and constants on an n x n grid (mass 1/N),
    positions perturbed by N(0, 1e-3*spacing), top edge pinned, no collider."""
    X = grid_positions(n, n, 4.0, 5.0)
    spacing = 4.0 / (n - 1)
    X = (X + (1e-3 * spacing) * normal(seed, X.size).reshape(X.shape)).astype(np.float32)
    dist, bend = pbd_grid_constraints(n, n, 0.25, 0.5)
    pinned = np.zeros(n * n, np.uint8)
    pinned[:n] = 1
    p = PBDParams(mass=float(np.float32(1.0) / np.float32(n * n)))
    return PBDScene(f"pbd_{n}x{n}", n, n, X, np.zeros_like(X), dist, bend, pinned, p, 100)
