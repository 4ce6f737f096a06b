"""GPU tests of the learned closure: the network's HIP forward/backward kernels (BF16 MFMA GEMMs) against the
CPU oracle of the same arithmetic, training on the GPU, the ADO extraction the training data comes from, the
learned correction of the terminator against the oracle, and what the committed trained network gains on
held-out parameter sets (4 spins, tiers 0..2 kept, reference = plain FP64 at tiers 0..6, 2000 steps)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NETWORK = os.path.join(ROOT, "helix_amd", "data", "closure_n16_L2.npz")


@pytest.fixture(scope="module")
def cl():
    from helix_amd import closure
    return closure


@pytest.mark.parametrize("dims,n", [((37, 32, 48, 24), 50), ((22, 256, 256, 2), 300)])
def test_mlp_forward_and_gradient_match_oracle(cl, dims, n):
    """Both operands of every GEMM are BF16 on the device and in the oracle; what remains is FP32 against FP64
    accumulation and tanh, which can move an activation across a BF16 rounding boundary now and then, so most
    outputs agree to 1e-5 and the rare moved ones to 2e-3."""
    from oracle import mlp_oracle as MO
    m = cl.MLP(*dims, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((n, dims[0])).astype(np.float32)
    t = rng.standard_normal((n, dims[3])).astype(np.float32)
    p = m.params()
    y = m.forward(x)
    y_ref, _ = MO.forward(p, x, dims, bf16=True)
    scale = max(1.0, np.max(np.abs(y_ref)))
    d = np.abs(y - y_ref) / scale
    print(f"forward: max {d.max():.2e}, share above 1e-5 {np.mean(d > 1e-5):.3f}")
    assert d.max() <= 2e-3 and np.mean(d > 1e-5) <= 0.05
    loss, g = m.gradient(x, t)
    loss_ref, g_ref = MO.loss_and_grad(p, x, t, dims, bf16=True)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    dg = np.abs(g - g_ref) / np.max(np.abs(g_ref))
    print(f"gradient: max {dg.max():.2e}, median {np.median(dg):.2e}")
    assert dg.max() <= 1e-2 and np.median(dg) <= 1e-5
    m.close()


def test_mlp_training_fits_a_smooth_map(cl):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (4096, 8)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] * x[:, 3] - 0.5 * x[:, 4]], 1).astype(np.float32)
    m = cl.MLP(8, 64, 64, 2, seed=9)
    losses = m.train(x, t, epochs=40, batch=128, lr=3e-3, seed=2)
    print("training losses", losses[0], losses[-1])
    assert losses[-1] < losses[0] / 20
    m.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_get_ados_matches_get_state(cl, precision):
    """helix_hip_get_ados / helix_hip_batch_get_ados return exactly the listed blocks of the full state."""
    import helix_amd as hx
    p = cl.Params.draw(9010)
    m = p.model(4, precision)
    for layout in ("auto", "full"):
        plan = hx.Plan(m, layout=layout)
        plan.set_initial()
        plan.step(30)
        full = plan.get_state()
        pick = np.array([0, 5, m.H - 1, 17, 5], np.int32)
        assert np.array_equal(plan.ados(pick), full[pick])
        plan.close()
    models = [cl.Params.draw(9011 + i).model(3 + i, precision) for i in range(3)]
    b = hx.Batch(models, graph=True, autotune=False)
    b.set_initial()
    b.step(20)
    for q, mq in enumerate(models):
        pick = np.arange(mq.H - 1, -1, -7, dtype=np.int32)
        assert np.array_equal(b.ados(q, pick), b.state(q)[pick])
    b.close()


def test_terminator_correction_matches_oracle(cl):
    """Right-hand side with a random correction table on the terminator against the oracle at 1e-12; a table of
    ones equals the physics terminator bit for bit; NULL restores it."""
    import helix_amd as hx
    from oracle import closure_oracle as CO
    p = cl.Params.draw(9012)
    term = cl.Terminator(p, 2, iterations=5)
    m = p.model(2, "fp64")
    rng = np.random.default_rng(3)
    x = rng.normal(size=(m.H, m.N, m.N)) + 1j * rng.normal(size=(m.H, m.N, m.N))
    corr = 1.0 + 0.3 * (rng.normal(size=(term.count, m.N, m.N)) + 1j * rng.normal(size=(term.count, m.N, m.N)))
    plan = hx.Plan(m, layout="full")
    cl.set_terminator(plan, term)
    plan.set_state(x)
    phys = plan.rhs()
    cl.set_terminator_gain(plan, corr)
    plan.set_state(x)
    got = plan.rhs()
    o = CO.spin_glass_terminator(4, p.seed, p.lam, p.gamma, p.T, 2)
    ref = o.rhs(x, 5, corr)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"corrected terminator rhs: normwise rel err {err:.2e}")
    assert err <= 1e-12
    cl.set_terminator_gain(plan, np.ones_like(corr))
    plan.set_state(x)
    assert np.array_equal(plan.rhs(), phys)
    cl.set_terminator_gain(plan, None)
    plan.set_state(x)
    assert np.array_equal(plan.rhs(), phys)
    plan.close()


def test_training_pipeline_on_a_few_sets(cl):
    """The whole pipeline at small size: response statistics from deep batched runs of 8 sets (tiers 0..4, 300
    steps), targets, a few epochs of training; the network then predicts the targets better than c = 1."""
    sets = [cl.Params.draw(7000 + i) for i in range(8)]
    terms, num, den = cl.response_statistics(sets, keep_tier=2, deep_tier=4, steps=300, every=50, chunk=4)
    X, Y = cl.correction_targets(terms, num, den, per_set=2048)
    assert X.shape[1] == cl.feature_dim(4) and len(X) > 1000
    m = cl.MLP(X.shape[1], 256, 256, 2, seed=11)
    losses = m.train(X, Y, epochs=30, batch=256, lr=1e-3, seed=3)
    base = float(np.mean(Y.astype(np.float64) ** 2))
    print(f"rows {len(X)}, target mean square {base:.3e}, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.5 * base
    m.close()


@pytest.fixture(scope="module")
def network(cl):
    if not os.path.exists(NETWORK):
        pytest.skip("trained network not present")
    m = cl.load_network(NETWORK)
    yield m
    m.close()


def test_learned_closure_gains_on_held_out_sets(cl, network):
    """Held-out parameter sets (seeds 9000..9005), 2000 complex64 steps (t = 20): max|d rho_S| over the trajectory
    against plain FP64 at tiers 0..6, for plain truncation at tier 2, the physics terminator and its learned
    correction."""
    import helix_amd as hx
    steps, every = 2000, 20
    errs = {"plain": [], "physics": [], "learned": []}
    for seed in range(9000, 9006):
        p = cl.Params.draw(seed)
        deep = hx.Plan(p.model(6, "fp64"))
        deep.set_initial()
        ref, _ = deep.step(steps, every, record_rho=True)
        deep.close()
        s = cl.ClosureSolver(p, network, keep_tier=2, precision="fp32")
        row = []
        for mode in ("plain", "physics", "learned"):
            s.use(mode)
            _, tr = s.run(steps, every)
            errs[mode].append(float(np.max(np.abs(tr - ref))))
            row.append(f"{mode} {errs[mode][-1]:.3e}")
        s.close()
        print(f"seed {seed}: " + "  ".join(row))
    mean = {k: float(np.mean(v)) for k, v in errs.items()}
    print("mean max|d rho_S|: " + "  ".join(f"{k} {v:.3e}" for k, v in mean.items()))
    assert mean["learned"] < 0.7 * mean["physics"] and mean["physics"] < mean["plain"]
    assert sum(a < b for a, b in zip(errs["learned"], errs["physics"])) >= 5


def test_closure_falls_back_to_plain_truncation(cl, network):
    """A calibration tolerance no run can meet forces the fallback; the result is then exactly the plain
    truncated run.  A generous tolerance keeps the closure."""
    p = cl.Params.draw(9100)
    s = cl.ClosureSolver(p, network, keep_tier=2, precision="fp32")
    assert s.calibrate("learned", window=100, tolerance=0.0) is False
    assert s.mode == "plain"
    fallback = s.run(500)
    s.use("plain")
    plain = s.run(500)
    assert np.array_equal(fallback, plain)
    assert s.calibrate("learned", window=100, tolerance=1.0) is True and s.mode == "learned"
    s.close()
