"""oc_cloth_step_noalloc: oc_cloth_step with the CG partials in the library's
fixed device array (no allocation in the call).  Two calls in a row, the
second on a smaller cloth, each bit-identical to oc_cloth_step on a twin
handle (state, last status, sticky status); the small one also against the
FP32 oracle.  A cloth whose partials do not fit the array is refused."""
import dataclasses

import numpy as np
import pytest

from opencloth_amd import scenes
from opencloth_amd._lib import OC_ERR_INVALID_ARGUMENT, check, lib
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu


def twin(sc):
    from opencloth_amd.cloth import ImplicitCloth

    return ImplicitCloth.from_scene(sc), ImplicitCloth.from_scene(sc)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_noalloc_step_equals_step_twice():
    from oracle import Oracle

    big = scenes.large_scene()
    small = scenes.plain_scene()
    for sc, n in ((big, 4), (small, 7)):
        say("noalloc step", sc.nx, "x", sc.ny, n, "steps")
        ref, cl = twin(sc)
        ref.step(n)
        check(lib().oc_cloth_step_noalloc(cl.h, n))
        Xr, Vr = ref.get_state()
        Xg, Vg = cl.get_state()
        assert cl.last_status() == ref.last_status()
        assert cl.sticky_status() == ref.sticky_status()
        assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(Vg), bits(Vr))
        if sc is small:
            s = Oracle(32).system(sc.n_vertices, sc.springs, sc.pinned)
            Xo = s.rollout(sc.X, sc.V, oracle_params(sc.params, sc.nx, sc.ny), n)[0]
            assert np.array_equal(bits(Xg), bits(Xo))
        ref.close()
        cl.close()


def test_noalloc_cold_start_and_refusal():
    """A cold CG start (an update launch per step) through the fixed array,
    then a cloth whose partials exceed it: OC_ERR_INVALID_ARGUMENT, state
    untouched."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    p = dataclasses.replace(sc.params, warm_start=False)
    ref = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    ref.step(3)
    check(lib().oc_cloth_step_noalloc(cl.h, 3))
    assert np.array_equal(bits(cl.get_state()[0]), bits(ref.get_state()[0]))
    ref.close()
    cl.close()

    big = scenes.large_scene()
    p = dataclasses.replace(big.params, cg_max_iters=2000)  # 512 tiles x 4001 floats > 2^20
    cl = ImplicitCloth(big.nx, big.ny, big.X, big.springs, big.pinned, p, big.V)
    rc = lib().oc_cloth_step_noalloc(cl.h, 1)
    assert rc == OC_ERR_INVALID_ARGUMENT
    assert b"fixed scratch" in lib().oc_last_error()
    X, _ = cl.get_state()
    assert np.array_equal(bits(X), bits(big.X))
    cl.close()
