"""GPU parity of the post-step passes (EllipsoidCollision IE:727-751,
ApplyProvotDynamicInverse IE:702-726), of the implicit demo's full
StepPhysics (IE:774-779) on its own scene, and of the position-Verlet step
(VL:428-562), all bit-exact against the FP32 oracle."""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_ellipsoid_bitwise(o32):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    ell = scenes.ellipsoid(translate=(0.0, 0.3, 2.0))
    rng = np.random.default_rng(1)
    X = (sc.X + 0.2 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V = rng.standard_normal(sc.X.shape).astype(np.float32)
    say("ellipsoid")
    cl = ImplicitCloth.from_scene(sc)
    cl.set_state(X, V)
    cl.ellipsoid_collision(ell)
    Xg, Vg = cl.get_state()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Xr, Vr = X.copy(), V.copy()
    s.ellipsoid(Xr, Vr, ell, 0)
    assert (Vr == 0).all(axis=1).sum() > 10  # some vertices were inside
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(Vg), bits(Vr))


@pytest.mark.parametrize("slack", [1.0, 1.01])
def test_provot_bitwise(o32, slack):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    rng = np.random.default_rng(2)
    X = (sc.X * 1.05 + 0.03 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V = rng.standard_normal(sc.X.shape).astype(np.float32)
    say("provot", slack)
    cl = ImplicitCloth.from_scene(sc)
    cl.set_state(X, V)
    cl.provot(slack)
    _, Vg = cl.get_state()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Vr = V.copy()
    s.provot(X.copy(), Vr, slack)
    assert not np.array_equal(Vr, V)
    assert np.array_equal(bits(Vg), bits(Vr))


def test_demo_step_physics_rollout(o32):
    """The implicit demo's scene and full StepPhysics (implicit step with
    ground clamp, ellipsoid, Provot) for 1300 steps (the cloth reaches the
    collider after about 1140), compared bit for bit every 50 steps."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.demo_scene()
    ell, slack = sc.extra["ellipsoid"], sc.extra["provot_slack"]
    say("demo StepPhysics x1300")
    cl = ImplicitCloth.from_scene(sc)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    X, V, DV = sc.X.copy(), sc.V.copy(), np.zeros_like(sc.X)
    for k in range(1300):
        s.step_physics(X, V, DV, prm, ell, slack)
        if k % 50 == 49:
            cl.step_physics(50, ell, slack)
            Xg, Vg = cl.get_state()
            assert np.array_equal(bits(Xg), bits(X)), k
            assert np.array_equal(bits(Vg), bits(V)), k
    assert X[:, 1].min() < 2.0  # passed the collider


def test_verlet_forces_bitwise(o32):
    from opencloth_amd.cloth import VerletCloth

    sc = scenes.verlet_demo_scene()
    rng = np.random.default_rng(3)
    X = (sc.X + 0.05 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    XL = (X - 0.01 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    p = sc.params
    vc = VerletCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, mass=p.mass, gravity=p.gravity,
                     damping=sc.extra["damping"], dt=p.dt, X_last=XL)
    F = vc.compute_forces()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Fr = s.verlet_forces(X, XL, p.dt, p.mass, p.gravity, sc.extra["damping"])
    assert np.array_equal(bits(F), bits(Fr))


def test_verlet_stencil_form_equals_slot_form(o32, monkeypatch):
    """A perturbed 64x48 grid (its spring list has the stencil assembly form,
    duplicated bends included): forces and 20 steps through the stencil form
    (springs from the ownership mask, constants from owner planes) equal the
    slot-table form (OPENCLOTH_ASSEMBLY=per-end) and the oracle bit for bit."""
    from opencloth_amd.cloth import VerletCloth

    nx, ny = 64, 48
    X0 = scenes.grid_positions(nx, ny)
    sp = scenes.grid_springs(nx, ny, X0, ks=(40.0, 30.0, 20.0), kd=(-0.2, -0.1, -0.05))
    rng = np.random.default_rng(5)
    X = (X0 + 0.02 * rng.standard_normal(X0.shape)).astype(np.float32)
    XL = (X - 0.005 * rng.standard_normal(X0.shape)).astype(np.float32)
    pinned = np.zeros(nx * ny, np.uint8)
    pinned[[0, nx - 1]] = 1
    kw = dict(mass=1.0, gravity=(0.0, -9.81, 0.0), damping=-0.0125, dt=1.0 / 60.0, X_last=XL)
    out = []
    for form in ("stencil", "per-end"):
        if form == "per-end":
            monkeypatch.setenv("OPENCLOTH_ASSEMBLY", "per-end")
        vc = VerletCloth(nx, ny, X, sp, pinned, **kw)
        F = vc.compute_forces()
        vc.step(20)
        out.append((F, vc.get_state()[0]))
        monkeypatch.delenv("OPENCLOTH_ASSEMBLY", raising=False)
    s = o32.system(nx * ny, sp, pinned)
    Fr = s.verlet_forces(X, XL, kw["dt"], kw["mass"], kw["gravity"], kw["damping"])
    assert np.array_equal(bits(out[0][0]), bits(Fr))
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][1]), bits(out[1][1]))


def test_verlet_demo_rollout(o32):
    """The Verlet demo's StepPhysics (forces, integrate, ground clamp,
    ellipsoid) for 300 steps, bit-exact."""
    from opencloth_amd.cloth import VerletCloth

    sc = scenes.verlet_demo_scene()
    p, ell = sc.params, sc.extra["ellipsoid"]
    say("verlet demo x300")
    vc = VerletCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, mass=p.mass, gravity=p.gravity,
                     damping=sc.extra["damping"], dt=p.dt, ground_clamp=True)
    vc.step(300, ell)
    Xg, XLg = vc.get_state()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Xr, XLr = s.verlet_rollout(sc.X, sc.X, p.dt, p.mass, p.gravity, sc.extra["damping"], 1, 300, ell)
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(XLg), bits(XLr))
    assert Xr[:, 1].min() < 4.9  # moved
