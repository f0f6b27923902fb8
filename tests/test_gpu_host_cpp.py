"""The C++ demo program (opencloth_amd/host/ie_demo.cpp: the implicit demo's
scene and StepPhysics loop through include/opencloth.hpp over the C ABI)
against the FP32 oracle on the same scene, bit for bit, after enough steps
for the cloth to reach the collider."""
import struct
import subprocess

import numpy as np
import pytest

from opencloth_amd import scenes
from opencloth_amd.build import host_program
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu


def test_cpp_demo_matches_oracle(tmp_path):
    from oracle import Oracle

    steps = 1200
    out = tmp_path / "state.bin"
    say("ie_demo", steps)
    r = subprocess.run([host_program("ie_demo"), "--steps", str(steps), "--out", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    say(r.stdout.strip())
    b = open(out, "rb").read()
    n, _ = struct.unpack("<ii", b[:8])
    Xg = np.frombuffer(b, np.float32, 3 * n, 8).reshape(n, 3)
    Vg = np.frombuffer(b, np.float32, 3 * n, 8 + 12 * n).reshape(n, 3)

    sc = scenes.demo_scene()
    ell, slack = sc.extra["ellipsoid"], sc.extra["provot_slack"]
    s = Oracle(32).system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    X, V, DV = sc.X.copy(), sc.V.copy(), np.zeros_like(sc.X)
    for _ in range(steps):
        s.step_physics(X, V, DV, prm, ell, slack)
    assert X[:, 1].min() < 2.5  # in contact with the collider
    assert np.array_equal(Xg.view(np.uint32), X.view(np.uint32))
    assert np.array_equal(Vg.view(np.uint32), V.view(np.uint32))
