"""GPU tests of the batched form: many independent problems of uneven N and depth in one launch per
Taylor stage, optionally replayed from a hipGraph.  Each element of the batched kernel runs the same
arithmetic as the single-problem kernel, so the batch must agree with single plans bit for bit; small
instances are also checked against the CPU FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


def _models(hx, seeds, precision="fp64"):
    return [hx.Model.spin_glass_random(s, dt=0.01, precision=precision) for s in seeds]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "split_fp32"])
def test_batch_equals_single_plans_bitwise(hx, precision):
    seeds = list(range(1000, 1012))
    models = _models(hx, seeds, precision)
    batch = hx.Batch(models, graph=True, autotune=True)
    batch.set_initial()
    batch.step(30)
    for q, s in enumerate(seeds):
        m = hx.Model.spin_glass_random(s, dt=0.01, precision=precision)
        plan = hx.Plan(m)
        plan.set_initial()
        plan.step(30)
        assert np.array_equal(batch.state(q), plan.get_state()), f"instance {q} (seed {s})"
        assert np.array_equal(batch.tier0(q), plan.tier0())
        plan.close()
    print(batch.info())


def test_batch_against_oracle(hx, oracle_mod):
    cases = [(3, 0, 2, (0.2, 0.8, 1.4)), (4, 1, 2, (0.05, 1.7, 0.6)), (3, 2, 3, (0.45, 1.1, 0.9))]
    models = [hx.Model.spin_glass(sp, sd, L=L, lam=p[0], gamma=p[1], T=p[2], dt=0.01) for sp, sd, L, p in cases]
    batch = hx.Batch(models)
    batch.set_initial()
    batch.step(100)
    for q, (sp, sd, L, p) in enumerate(cases):
        o = oracle_mod.spin_glass_problem(sp, sd, L=L, lam=p[0], gamma=p[1], T=p[2])
        ref = o.taylor(o.initial(), 0.01, 100)
        got = batch.state(q)
        assert float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))) <= FP64_TOL


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_ragged_with_depth_zero_and_one(hx, oracle_mod, persistent):
    """A ragged batch: a one-ADO instance (tiers 0..0), a tier-1 one, and deeper ones of every block size
    (N = 8, 16, 32, 64), per-size and persistent launches; each instance against the oracle over 50 steps."""
    cases = [(3, 0, 0), (6, 1, 0), (4, 2, 1), (5, 3, 3), (6, 4, 2), (3, 5, 4)]
    models = [hx.Model.spin_glass(sp, sd, L=L, dt=0.01) for sp, sd, L in cases]
    batch = hx.Batch(models, graph=True, autotune=False, persistent=persistent)
    batch.set_initial()
    batch.step(50)
    for q, (sp, sd, L) in enumerate(cases):
        o = oracle_mod.spin_glass_problem(sp, sd, L=L)
        ref = o.taylor(o.initial(), 0.01, 50)
        got = batch.state(q)
        assert got.shape == ref.shape
        assert float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))) <= FP64_TOL, f"instance {q}: {sp} spins, L={L}"
    batch.close()


def test_batch_graph_replays_match_eager(hx):
    seeds = list(range(1020, 1028))
    g = hx.Batch(_models(hx, seeds), graph=True, autotune=False)
    e = hx.Batch(_models(hx, seeds), graph=False, autotune=False)
    assert g.info()["graph"] and not e.info()["graph"]
    g.set_initial()
    e.set_initial()
    for rnd in range(5):
        g.step(3)
        e.step(3)
        for q in range(len(seeds)):
            assert np.array_equal(g.state(q), e.state(q)), f"replay round {rnd}, instance {q}"


def test_batch_complex64_form_error(hx):
    """Batched complex64 against batched FP64 over 500 steps: max|d rho_S| <= 1e-6 per instance."""
    seeds = list(range(1030, 1046))
    b64 = hx.Batch(_models(hx, seeds, "fp64"))
    b32 = hx.Batch(_models(hx, seeds, "fp32"))
    b64.set_initial()
    b32.set_initial()
    b64.step(500)
    b32.step(500)
    worst = max(float(np.max(np.abs(b32.tier0(q) - b64.tier0(q)))) for q in range(len(seeds)))
    print(f"batched complex64: worst max|d rho_S| over {len(seeds)} instances = {worst:.3e}")
    assert worst <= 1e-6


def test_batch_rejects_bit_flip_couplings(hx):
    with pytest.raises(hx.HelixError):
        hx.Batch([hx.Model.legacy(N=32)])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_persistent_launch_equals_per_size_launches(hx, precision):
    """One persistent launch per stage (per-XCD task queues) and one launch per block size run the same tile bodies:
    bit-identical states, and equal to single plans; the persistent form makes one launch per stage."""
    seeds = list(range(1000, 1024))
    runs = {}
    for per_size in (False, True):
        b = hx.Batch(_models(hx, seeds, precision), graph=True, autotune=False, per_size=per_size, persistent=not per_size)
        b.set_initial()
        b.step(12)
        info = b.info()
        assert info["launches_per_step"] == (4 * len({m.N for m in b.models}) if per_size else 4), info
        runs[per_size] = [b.state(q) for q in range(len(seeds))]
        b.close()
    for q in range(len(seeds)):
        assert np.array_equal(runs[False][q], runs[True][q]), f"instance {q}"
    for q in (0, 7, 23):
        plan = hx.Plan(hx.Model.spin_glass_random(seeds[q], dt=0.01, precision=precision))
        plan.set_initial()
        plan.step(12)
        assert np.array_equal(runs[False][q], plan.get_state())
        plan.close()
