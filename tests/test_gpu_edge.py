"""Edge cases of the implicit step on the GPU, each against the FP32 oracle
with the kernels' reduction order (positions, velocities, iteration count
and status bits): the smallest grid, grids narrower or shorter than a tile,
every vertex pinned, a spring list that is a subset of the reference's
(only structural springs; a random subset) or a permutation of it, a NaN in the state (the
non-finite flag), and a compressed stiff cloth whose system is indefinite
(the breakdown flag)."""
import dataclasses

import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu

NONFINITE, BREAKDOWN = 4, 2


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(a, b):
    """Bit-equal, except that a NaN matches any NaN (payloads are not part
    of the contract)."""
    a = np.asarray(a, np.float32); b = np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(bits(a[~na]), bits(b[~nb]))


def compare(o32, nx, ny, X, V, springs, pinned, p, nsteps, expect_status=None):
    from opencloth_amd.cloth import ImplicitCloth

    cl = ImplicitCloth(nx, ny, X, springs, pinned, p, V)
    s = o32.system(nx * ny, springs, pinned)
    prm = oracle_params(p, nx, ny)
    Xr, Vr = X.astype(np.float32).copy(), V.astype(np.float32).copy()
    DV = np.zeros_like(Xr)
    for k in range(nsteps):
        it_r, st_r = s.step(Xr, Vr, DV, prm)
        cl.step(1)
        it_g, st_g = cl.last_status()
        assert (it_g, st_g) == (it_r, st_r), (k, it_g, st_g, it_r, st_r)
        if expect_status is not None:
            assert st_g & expect_status
    Xg, Vg = cl.get_state()
    assert same(Xg, Xr) and same(Vg, Vr)
    return Xg, Vg, it_r, st_r


def springs_2x2(X0):
    """The 2x2 grid's four edges and two diagonals (the reference's list
    needs 3x3 for its bend springs)."""
    s = np.zeros(6, scenes.SPRING_DTYPE)
    s["p1"] = [0, 2, 0, 1, 0, 2]
    s["p2"] = [1, 3, 2, 3, 3, 1]
    s["type"] = [0, 0, 0, 0, 1, 1]
    s["rest_length"] = scenes.rest_lengths(X0, s["p1"], s["p2"])
    s["Ks"], s["Kd"] = 50.0, -0.25
    return s


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 70), (70, 3), (65, 9), (5, 5)])
def test_grid_shapes(o32, nx, ny):
    """Smallest grid, one-tile-wide, one-tile-high and tile-straddling grids:
    4 steps bit-exact (CG up to 10 iterations, absolute tolerance 1e-7)."""
    X0 = scenes.grid_positions(nx, ny, 4.0, 0.0)
    X = scenes.perturb(X0, 0.01, nx * 100 + ny)
    springs = springs_2x2(X0) if nx == 2 else scenes.grid_springs(nx, ny, X0)
    pinned = np.zeros(nx * ny, np.uint8)
    pinned[0] = 1
    V = np.zeros_like(X)
    say("grid", nx, ny)
    compare(o32, nx, ny, X, V, springs, pinned, scenes.StepParams(dt=1 / 60, cg_tol=1e-7), 4)


def test_all_pinned(o32):
    """Every vertex pinned: r = 0 from the start, no CG iteration, nothing moves."""
    sc = scenes.plain_scene()
    pinned = np.ones(sc.n_vertices, np.uint8)
    V = np.zeros_like(sc.X)
    Xg, Vg, it, st = compare(o32, sc.nx, sc.ny, sc.X, V, sc.springs, pinned, sc.params, 2)
    assert it == 0 and st == 0
    assert same(Xg, sc.X) and not Vg.any()


@pytest.mark.parametrize("which", ["structural", "random"])
def test_spring_subsets(o32, which):
    """Spring lists other than the reference's full list."""
    sc = scenes.plain_scene()
    sp = sc.springs
    if which == "structural":
        sp = sp[sp["type"] == scenes.STRUCTURAL]
    else:
        keep = np.random.default_rng(7).random(len(sp)) < 0.6
        sp = sp[keep]
    say("subset", which, len(sp))
    X = scenes.perturb(sc.X, 0.02, 3)
    p = dataclasses.replace(sc.params, cg_tol=1e-7)
    compare(o32, sc.nx, sc.ny, X, np.zeros_like(X), np.ascontiguousarray(sp), sc.pinned, p, 5)


def test_shuffled_spring_list(o32):
    """The reference's springs in a random order: pairs are numbered by their
    first spring, so every vertex sums its pairs in another order (the
    kernels' general slot tables instead of the stencil mask)."""
    sc = scenes.plain_scene()
    perm = np.random.default_rng(5).permutation(len(sc.springs))
    sp = np.ascontiguousarray(sc.springs[perm])
    X = scenes.perturb(sc.X, 0.02, 9)
    p = dataclasses.replace(sc.params, cg_tol=1e-7)
    compare(o32, sc.nx, sc.ny, X, np.zeros_like(X), sp, sc.pinned, p, 5)


def test_springs_off_the_stencil(o32):
    """Springs two columns and one row apart (outside the reference's 12
    neighbours, inside the 2-wide halo) appended to the reference list:
    the general slot tables, bit-exact."""
    sc = scenes.plain_scene()
    n = sc.nx
    extra = np.zeros(30, scenes.SPRING_DTYPE)
    rows = np.arange(30) % (n - 1)
    cols = (np.arange(30) * 7) % (n - 2)
    extra["p1"] = rows * n + cols
    extra["p2"] = (rows + 1) * n + cols + 2
    extra["rest_length"] = scenes.rest_lengths(sc.X, extra["p1"], extra["p2"])
    extra["Ks"], extra["Kd"], extra["type"] = 30.0, -0.25, 2
    sp = np.concatenate([sc.springs, extra])
    X = scenes.perturb(sc.X, 0.02, 17)
    p = dataclasses.replace(sc.params, cg_tol=1e-7)
    compare(o32, sc.nx, sc.ny, X, np.zeros_like(X), sp, sc.pinned, p, 4)


def test_no_springs_free_fall(o32):
    """No springs at all: gravity only (free vertices fall, pins stay)."""
    sc = scenes.plain_scene()
    sp = np.zeros(0, scenes.SPRING_DTYPE)
    Xg, _, _, _ = compare(o32, sc.nx, sc.ny, sc.X, np.zeros_like(sc.X), sp, sc.pinned, sc.params, 3)
    assert np.array_equal(Xg[sc.pinned == 1], sc.X[sc.pinned == 1])
    assert (Xg[sc.pinned == 0, 1] < sc.X[sc.pinned == 0, 1]).all()


def test_batch_with_a_springless_cloth(o32):
    """A batch whose middle cloth has no springs steps every cloth bit-exactly
    against the oracle."""
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sides = [(12, 9), (7, 5), (10, 10)]
    Xs, sps, pins, ns = [], [], [], []
    for c, (nx, ny) in enumerate(sides):
        X0 = scenes.grid_positions(nx, ny, 4.0, 0.0)
        sp = scenes.grid_springs(nx, ny, X0) if c != 1 else np.zeros(0, scenes.SPRING_DTYPE)
        pin = np.zeros(nx * ny, np.uint8)
        pin[:2] = 1
        Xs.append(scenes.perturb(X0, 0.01, c)); sps.append(sp); pins.append(pin); ns.append(len(sp))
    p = scenes.StepParams(dt=1 / 60, cg_max_iters=20, cg_tol=1e-6)
    b = ClothBatch(np.array([a for a, _ in sides]), np.array([b_ for _, b_ in sides]), np.concatenate(Xs),
                   np.concatenate(sps), np.array(ns), np.concatenate(pins), p)
    b.set_schedule(256, 4)
    T = b.threads_per_cloth()
    b.step(3, graph=False)
    X, _ = b.get_state()
    off = 0
    for c, (nx, ny) in enumerate(sides):
        s = o32.system(nx * ny, sps[c], pins[c])
        Xr = s.rollout(Xs[c], np.zeros_like(Xs[c]), make_params(p, dot_mode=2, tx=int(T[c])), 3)[0]
        assert same(X[off:off + nx * ny], Xr), c
        off += nx * ny


def test_nan_state_sets_nonfinite(o32):
    """A NaN position: the status carries OC_SOLVE_NONFINITE, as in the
    oracle, and the same entries are NaN afterwards."""
    sc = scenes.plain_scene()
    X = sc.X.copy()
    X[200, 1] = np.nan
    compare(o32, sc.nx, sc.ny, X, np.zeros_like(X), sc.springs, sc.pinned, sc.params, 1,
            expect_status=NONFINITE)


def test_indefinite_system_reports_breakdown(o32):
    """Compressed to 40% with stiff springs and a long step: A is indefinite
    (the reference applies no clamp), CG meets d.q < 0 and both report the
    breakdown at the same iteration."""
    sc = scenes.plain_scene()
    X = (sc.X * 0.4).astype(np.float32)
    springs = sc.springs.copy()
    springs["Ks"] = 5000.0
    p = dataclasses.replace(sc.params, dt=0.1, cg_max_iters=40, cg_tol=1e-9)
    Xg, Vg, it, st = compare(o32, sc.nx, sc.ny, X, np.zeros_like(X), springs, sc.pinned, p, 1,
                             expect_status=BREAKDOWN)
    say("breakdown at iteration", it)
