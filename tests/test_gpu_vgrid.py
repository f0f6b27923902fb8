"""GPU parity of the grid Verlet (oc_vgrid_*, the GPU demos' kernel
verlet_kernel.cu:70-197) through the C ABI: bit for bit against the FP32
oracle (same expression order, no contraction) on the demo's 64x64 cloth
and on ragged grids with random pins, previous states, vertices below the
ground and non-square spacing; the default pins; the vertex-buffer output;
parameters changed between steps; state round trips with w."""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_demo_scene_bitwise(o32):
    from opencloth_amd.cloth import VerletGrid

    sc = scenes.vgrid_demo_scene(64)
    say("vgrid demo 64x64, 300 steps")
    g = VerletGrid.from_scene(sc)
    g.step(300)
    X, XL = g.get_state()
    Xr, XLr = o32.vgrid_rollout(sc.width, sc.height, sc.X, sc.X_last, sc.pinned, sc.params, 300)
    assert np.array_equal(bits(X), bits(Xr)) and np.array_equal(bits(XL), bits(XLr))
    assert np.abs(X[:, :3] - sc.X[:, :3]).max() > 0.1


@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 5), (3, 7), (33, 9), (130, 70), (40, 33)])
def test_ragged_grids_bitwise(o32, w, h):
    from opencloth_amd.cloth import VerletGrid

    rng = np.random.default_rng(w * 131 + h)
    p = scenes.vgrid_params(w + 1, h + 3)
    p["damping"], p["gravity"] = -0.05, (0.01, -0.3, 0.02)
    sx, sy = p["spacing"]
    ix, iy = np.meshgrid(np.arange(w), np.arange(h))
    X = np.stack([ix.ravel() * sx, rng.uniform(-0.05, 0.05, w * h), iy.ravel() * sy,
                  rng.random(w * h)], 1).astype(np.float32)  # about half below the ground
    X[:, :3] += (0.2 * min(sx, sy) * rng.standard_normal((w * h, 3))).astype(np.float32)
    XL = X.copy()
    XL[:, :3] -= (0.01 * rng.standard_normal((w * h, 3))).astype(np.float32)
    XL[:, 3] = -3.0
    pin = (rng.random(w * h) < 0.2).astype(np.uint8)
    say("vgrid", w, "x", h)
    g = VerletGrid(w, h, X, XL, pin, p)
    g.step(6)
    Xg, XLg = g.get_state()
    Xr, XLr = o32.vgrid_rollout(w, h, X, XL, pin, p, 6)
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(XLg), bits(XLr))


def test_default_pins_and_vbo(o32):
    from opencloth_amd.cloth import VerletGrid
    from opencloth_amd.hiptime import DeviceBuffer

    sc = scenes.vgrid_large_scene(96, seed=4)
    a = VerletGrid(sc.width, sc.height, sc.X, sc.X_last, None, sc.params)  # pins 0 and width-1
    b = VerletGrid.from_scene(sc)
    vbo = DeviceBuffer(16 * sc.width * sc.height)
    vbo.fill_zero()
    a.step(4, vbo=vbo.ptr)
    b.step(4)
    Xa, XLa = a.get_state()
    Xb, XLb = b.get_state()
    assert np.array_equal(bits(Xa), bits(Xb)) and np.array_equal(bits(XLa), bits(XLb))
    out = np.zeros((sc.width * sc.height, 4), np.float32)
    vbo.to_host(out)
    assert np.array_equal(bits(out), bits(Xa))  # last step's positions, w carried


def test_params_between_steps_and_round_trip(o32):
    from opencloth_amd.cloth import VerletGrid

    sc = scenes.vgrid_demo_scene(40)
    g = VerletGrid.from_scene(sc)
    p2 = dict(sc.params, dt=1.0 / 90.0, ks=(30.0, 20.0, 10.0), damping=-0.02)
    g.step(5)
    g.set_params(p2)
    g.step(5)
    X, XL = g.get_state()
    Xr, XLr = o32.vgrid_rollout(sc.width, sc.height, sc.X, sc.X_last, sc.pinned, sc.params, 5)
    Xr, XLr = o32.vgrid_rollout(sc.width, sc.height, Xr, XLr, sc.pinned, p2, 5)
    assert np.array_equal(bits(X), bits(Xr)) and np.array_equal(bits(XL), bits(XLr))
    rng = np.random.default_rng(2)
    X2 = (X + 0.01 * rng.standard_normal(X.shape)).astype(np.float32)
    g.set_state(X=X2)
    Xg, XLg = g.get_state()
    assert np.array_equal(Xg, X2) and np.array_equal(XLg, XL)


def test_caller_stream_and_rest_start(o32):
    """Work ordered on a caller's HIP stream gives the same bits, and a NULL
    x_last starts at rest (x_last = x)."""
    from opencloth_amd.cloth import VerletGrid
    from opencloth_amd.hiptime import Stream

    sc = scenes.vgrid_large_scene(70, seed=9)
    ref = VerletGrid(sc.width, sc.height, sc.X, sc.X, sc.pinned, sc.params)
    ref.step(7)
    g = VerletGrid(sc.width, sc.height, sc.X, None, sc.pinned, sc.params)
    stream = Stream()
    g.set_stream(stream.handle)
    g.step(7)
    g.synchronize()
    Xg, XLg = g.get_state()
    g.set_stream(None)
    Xr, XLr = ref.get_state()
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(XLg), bits(XLr))
    Xo, XLo = o32.vgrid_rollout(sc.width, sc.height, sc.X, sc.X, sc.pinned, sc.params, 7)
    assert np.array_equal(bits(Xg), bits(Xo)) and np.array_equal(bits(XLg), bits(XLo))
