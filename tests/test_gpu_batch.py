"""GPU parity of the batched step (oc_batch_*) through the C ABI.

Bars: every cloth of a batch bit-exact against the FP32 oracle run of that
cloth alone with the batched kernel's reduction order (threads per cloth from
the schedule); graph replay bit-exact against direct launches; the batch
within 1e-5 relative of the single-cloth GPU path (different reduction
order); autotuning leaves the state untouched."""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import rel_pos_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def small_batch(n=14, seed=3, sides=(16, 40)):
    return scenes.batch_scene(n_cloths=n, sides=sides, seed=seed, steps=4)


@pytest.mark.parametrize("sched", [(256, 4), (512, 8), (1024, 64), (256, 1000)])
def test_batch_bitwise_against_oracle(o32, sched):
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = small_batch()
    say("batch", sc.n_cloths, "cloths, schedule", sched)
    b = ClothBatch.from_scene(sc)
    b.set_schedule(*sched)
    T = b.threads_per_cloth()
    nsteps = 3
    b.step(nsteps, graph=False)
    X, V = b.get_state()
    it, st = b.status()
    voff = sc.voff
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        s = o32.system(cs.n_vertices, cs.springs, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        Xr, Vr, _, _, iters, str_ = s.rollout(cs.X, cs.V, prm, nsteps)
        n = cs.n_vertices
        Xg = X[voff[c]:voff[c] + n]
        assert it[c] == iters[-1] and st[c] == str_, (c, it[c], iters, st[c], str_)
        assert np.array_equal(Xg.view(np.uint32), Xr.view(np.uint32)), (c, np.abs(Xg - Xr).max())


@pytest.mark.parametrize("sched", [(1024, 1), (256, 1)])
def test_batch_large_cloths_bitwise(o32, sched):
    """Cloths of 200-260 vertices a side: every thread serves 40 to 260
    vertices, the first three with their CG vectors in LDS and the rest in
    memory; bit-exact against the oracle."""
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = scenes.batch_scene(n_cloths=3, sides=(200, 260), seed=11, steps=3)
    say("batch of large cloths", list(zip(sc.nx.tolist(), sc.ny.tolist())), "schedule", sched)
    b = ClothBatch.from_scene(sc)
    b.set_schedule(*sched)
    T = b.threads_per_cloth()
    b.step(3, graph=False)
    X, V = b.get_state()
    it, st = b.status()
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        s = o32.system(cs.n_vertices, cs.springs, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        Xr, Vr, _, _, iters, str_ = s.rollout(cs.X, cs.V, prm, 3)
        n = cs.n_vertices
        assert it[c] == iters[-1] and st[c] == str_, (c, it[c], iters, st[c], str_)
        assert np.array_equal(X[sc.voff[c]:sc.voff[c] + n].view(np.uint32), Xr.view(np.uint32)), c
        assert np.array_equal(V[sc.voff[c]:sc.voff[c] + n].view(np.uint32), Vr.view(np.uint32)), c


def test_batch_general_slot_tables(o32):
    """One cloth's springs in a shuffled order: that cloth sums its pairs in
    another order, so the whole batch runs on the general slot tables
    instead of the stencil mask; bit-exact against the oracle."""
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=5, seed=13, sides=(16, 30))
    n0 = int(sc.n_springs[0])
    perm = np.random.default_rng(4).permutation(n0)
    sc.springs[:n0] = sc.springs[:n0][perm]
    b = ClothBatch.from_scene(sc)
    b.set_schedule(256, 4)
    T = b.threads_per_cloth()
    b.step(3, graph=False)
    X, _ = b.get_state()
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        s = o32.system(cs.n_vertices, cs.springs, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        Xr = s.rollout(cs.X, cs.V, prm, 3)[0]
        Xg = X[sc.voff[c]:sc.voff[c] + cs.n_vertices]
        assert np.array_equal(Xg.view(np.uint32), Xr.view(np.uint32)), c


def test_batch_rejected_schedule_and_partial_state():
    """A rejected schedule leaves the previous one in force (same threads
    per cloth, a captured step still replays bit-exactly); set_state with
    one of x and v leaves the other unchanged."""
    from opencloth_amd._lib import OpenClothError
    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=6, seed=5, sides=(16, 32))
    a = ClothBatch.from_scene(sc)
    b = ClothBatch.from_scene(sc)
    a.set_schedule(512, 4)
    b.set_schedule(512, 4)
    T = a.threads_per_cloth()
    a.step(1, graph=True)  # captures a step with the 512/4 tables
    b.step(1, graph=True)
    for bad in ((384, 4), (512, 0)):
        with pytest.raises(OpenClothError):
            a.set_schedule(*bad)
        assert np.array_equal(a.threads_per_cloth(), T)
    a.step(2, graph=True)
    b.step(2, graph=True)
    Xa, Va = a.get_state()
    Xb, Vb = b.get_state()
    assert np.array_equal(Xa.view(np.uint32), Xb.view(np.uint32))
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32))
    a.set_state(V=np.zeros_like(Va))
    X1, V1 = a.get_state()
    assert np.array_equal(X1, Xa) and not V1.any()
    a.set_state(X=Xa + 1.0)
    X2, V2 = a.get_state()
    assert np.array_equal(X2, Xa + 1.0) and not V2.any()
    a.close()
    b.close()


def test_batch_spring_stiffness_update(o32):
    """Per-spring stiffness set after creation (the kernels read a slot-major
    copy of it) round-trips and steps every cloth bit-exactly against the
    oracle with those stiffnesses."""
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=6, seed=11, sides=(16, 36))
    b = ClothBatch.from_scene(sc)
    b.set_schedule(512, 4)
    S = len(sc.springs)
    ks = np.exp(np.random.default_rng(2).uniform(np.log(20.0), np.log(200.0), S)).astype(np.float32)
    b.set_spring_stiffness(ks)
    assert np.array_equal(b.spring_stiffness(S), ks)
    T = b.threads_per_cloth()
    b.step(3, graph=True)
    X, _ = b.get_state()
    soff = np.concatenate([[0], np.cumsum(sc.n_springs)])
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        sp = cs.springs.copy()
        sp["Ks"] = ks[soff[c]:soff[c + 1]]
        s = o32.system(cs.n_vertices, sp, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        Xr = s.rollout(cs.X, cs.V, prm, 3)[0]
        Xg = X[sc.voff[c]:sc.voff[c] + cs.n_vertices]
        assert np.array_equal(Xg.view(np.uint32), Xr.view(np.uint32)), c


def test_batch_graph_equals_eager():
    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=20, seed=5, sides=(16, 64))
    say("graph vs eager")
    a = ClothBatch.from_scene(sc)
    b = ClothBatch.from_scene(sc)
    a.step(5, graph=True)
    for _ in range(5):
        b.step(1, graph=False)
        b.synchronize()
    Xa, Va = a.get_state()
    Xb, Vb = b.get_state()
    assert np.array_equal(Xa.view(np.uint32), Xb.view(np.uint32))
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32))


@pytest.mark.parametrize("sched", [(256, 1), (512, 1), (1024, 1), (512, 32)])
def test_batch_spring_once_form_bitwise(monkeypatch, sched):
    """The opt-in spring-once batched assembly (OPENCLOTH_ASSEMBLY=once: each
    pair evaluated by its lower vertex, the partner reading the result from
    LDS, owners of earlier bands from rings) equals the default form bit for
    bit: 3 steps and the state, on cloths up to 128 wide, including
    schedules whose threads per cloth are fewer than two grid rows."""
    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=12, seed=9, sides=(16, 128))
    say("spring-once batch form, schedule", sched)
    ref = ClothBatch.from_scene(sc)
    monkeypatch.setenv("OPENCLOTH_ASSEMBLY", "once")
    once = ClothBatch.from_scene(sc)
    monkeypatch.delenv("OPENCLOTH_ASSEMBLY")
    for b in (ref, once):
        b.set_schedule(*sched)
        b.step(3, graph=False)
    Xa, Va = ref.get_state()
    Xb, Vb = once.get_state()
    ita, sta = ref.status()
    itb, stb = once.status()
    ref.close()
    once.close()
    assert np.array_equal(ita, itb) and np.array_equal(sta, stb)
    assert np.array_equal(Xa.view(np.uint32), Xb.view(np.uint32))
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32))


def test_batch_close_to_single_cloth_path():
    from opencloth_amd.batch import ClothBatch
    from opencloth_amd.cloth import ImplicitCloth

    sc = small_batch(n=6, seed=7, sides=(20, 48))
    say("batch vs single")
    b = ClothBatch.from_scene(sc)
    b.step(5)
    X, _ = b.get_state()
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        cl = ImplicitCloth.from_scene(cs)
        cl.step(5)
        n = cs.n_vertices
        err = rel_pos_err(X[sc.voff[c]:sc.voff[c] + n], cl.X)
        assert err < 1e-5, (c, err)


def test_autotune_keeps_state_and_picks_a_schedule():
    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=40, seed=9, sides=(16, 96))
    say("autotune")
    b = ClothBatch.from_scene(sc)
    b.step(2)
    X0, V0 = b.get_state()
    best = b.autotune()
    X1, V1 = b.get_state()
    assert best[0] in (256, 512, 1024) and best[1] >= 1
    assert b.schedule() == best
    assert np.array_equal(X0, X1) and np.array_equal(V0, V1)
    b.step(1)
    it, st = b.status()
    assert np.all(st == 0) and np.all(it <= sc.params.cg_max_iters)


def test_batch_deterministic():
    from opencloth_amd.batch import ClothBatch

    sc = small_batch(n=30, seed=11, sides=(16, 80))
    xs = []
    for _ in range(2):
        b = ClothBatch.from_scene(sc)
        b.set_schedule(512, 4)
        b.step(4)
        xs.append(b.get_state()[0])
    assert np.array_equal(xs[0].view(np.uint32), xs[1].view(np.uint32))
