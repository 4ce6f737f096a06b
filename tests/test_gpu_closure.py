"""GPU tests of the learned closure: the HIP forward/backward kernels against the CPU float64 oracle
(with the same BF16-rounded weights), training on the GPU, and what the trained closure gains over
plain truncation on held-out parameter sets (4 spins, tiers 0..2 kept, reference = plain FP64 at 0..6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cl():
    from helix_amd import closure
    return closure


def test_mlp_forward_and_gradient_match_oracle(cl):
    from oracle import mlp_oracle as MO
    dims = (37, 32, 48, 24)
    m = cl.MLP(*dims, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((50, dims[0])).astype(np.float32)
    t = rng.standard_normal((50, dims[3])).astype(np.float32)
    p = m.params()
    y = m.forward(x)
    y_ref, _ = MO.forward(p, x, dims, bf16=True)
    assert np.max(np.abs(y - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = m.gradient(x, t)
    loss_ref, g_ref = MO.loss_and_grad(p, x, t, dims, bf16=True)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * np.max(np.abs(g_ref))
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


@pytest.fixture(scope="module")
def trained(cl):
    train_sets = [cl.Params.draw(7000 + i) for i in range(128)]
    X, Y = cl.training_pairs(train_sets, keep_tier=2, deep_tier=6, steps=2000, every=100)
    m = cl.MLP(X.shape[1], 256, 256, Y.shape[1], seed=11)
    losses = m.train(X, Y, epochs=20, batch=512, lr=1e-3, seed=3)
    print(f"closure data {X.shape}, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    yield m
    m.close()


def test_learned_closure_beats_plain_truncation(cl, trained):
    """Held-out parameter sets, 2000 steps (t = 20): error of rho_S against plain FP64 at tiers 0..6 for
    plain truncation at tier 2, the learned closure, and the closure behind its calibration fallback
    (500-step window against the deep plain run)."""
    import helix_amd as hx
    steps = 2000
    errs_plain, errs_learned, errs_guarded = [], [], []
    for seed in (9000, 9001, 9002, 9003, 9004, 9005):
        p = cl.Params.draw(seed)
        deep = hx.Plan(p.model(6, "fp64"))
        deep.set_initial()
        deep.step(steps)
        ref = deep.tier0()
        deep.close()
        s = cl.LearnedClosureSolver(p, trained, keep_tier=2, precision="fp32")
        s.enable(False)
        plain = s.run(steps)
        s.enable(True)
        learned = s.run(steps)
        trusted = s.calibrate(window=500)
        guarded = s.run(steps)
        s.close()
        errs_plain.append(float(np.max(np.abs(plain - ref))))
        errs_learned.append(float(np.max(np.abs(learned - ref))))
        errs_guarded.append(float(np.max(np.abs(guarded - ref))))
        print(f"seed {seed}: plain L=2 {errs_plain[-1]:.3e}  learned {errs_learned[-1]:.3e}  "
              f"with fallback {errs_guarded[-1]:.3e} (closure {'kept' if trusted else 'dropped'})")
    gain = np.mean(errs_plain) / np.mean(errs_learned)
    gain_g = np.mean(errs_plain) / np.mean(errs_guarded)
    print(f"mean error plain {np.mean(errs_plain):.3e} learned {np.mean(errs_learned):.3e} ({gain:.2f}x) "
          f"with fallback {np.mean(errs_guarded):.3e} ({gain_g:.2f}x)")
    assert np.mean(errs_learned) < np.mean(errs_plain)
    assert np.mean(errs_guarded) < np.mean(errs_plain)


def test_closure_falls_back_to_plain_truncation(cl, trained):
    """A calibration tolerance no run can meet forces the fallback; the result is then exactly the plain
    truncated run."""
    p = cl.Params.draw(9100)
    s = cl.LearnedClosureSolver(p, trained, keep_tier=2, precision="fp32")
    assert s.calibrate(window=100, tolerance=0.0) is False
    assert s.closure_on is False
    fallback = s.run(500)
    s.enable(False)
    plain = s.run(500)
    assert np.array_equal(fallback, plain)
    s.close()
