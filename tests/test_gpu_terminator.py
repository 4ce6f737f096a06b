"""GPU tests of the physics terminator (helix_hip_plan_set_terminator) against the CPU oracle
(oracle/closure_oracle.py): the right-hand side and trajectories of a hierarchy kept to tiers 0..2 with the
tier-3 ADOs solved in every stage, FP64 at the plain-form bar and the FP32 form at 1e-6; switching it off
restores plain truncation bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITER = 5


@pytest.fixture(scope="module")
def cl():
    from helix_amd import closure
    return closure


def _plan(cl, p, precision="fp64", keep=2):
    import helix_amd as hx
    m = p.model(keep, precision)
    return hx.Plan(m, layout="full"), m


def _oracle(p, keep=2):
    from oracle import closure_oracle as CO
    return CO.spin_glass_terminator(4, p.seed, p.lam, p.gamma, p.T, keep)


@pytest.mark.parametrize("seed", [9000, 9003])
def test_terminator_rhs_matches_oracle(cl, seed):
    p = cl.Params.draw(seed)
    print(f"terminator rhs: seed {seed}", flush=True)
    plan, m = _plan(cl, p)
    cl.set_terminator(plan, cl.Terminator(p, 2, iterations=ITER))
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(m.H, m.N, m.N)) + 1j * rng.normal(size=(m.H, m.N, m.N))
    plan.set_state(x)
    got = plan.rhs()
    ref = _oracle(p).rhs(x, ITER)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"normwise rel err {err:.2e}")
    assert err <= 1e-12
    plan.close()


def test_terminator_trajectory_fp64_and_fp32(cl):
    p = cl.Params.draw(9001)
    print("terminator trajectory: seed 9001, 300 steps", flush=True)
    o = _oracle(p)
    ref = o.taylor(o.small.initial(), 0.01, 300, iterations=ITER)
    for prec in ("fp64", "fp32"):
        plan, m = _plan(cl, p, prec)
        cl.set_terminator(plan, cl.Terminator(p, 2, iterations=ITER))
        plan.set_initial()
        plan.step(300)
        if prec == "fp64":
            err = float(np.max(np.abs(plan.get_state() - ref)) / np.max(np.abs(ref)))
            assert err <= 1e-12, err
        else:
            err = float(np.max(np.abs(plan.tier0() - ref[0])))
            print(f"fp32 max|d rho_S| = {err:.2e}")
            assert err <= 1e-6
        plan.close()


def test_terminator_off_restores_plain_truncation(cl):
    p = cl.Params.draw(9002)
    plain, _ = _plan(cl, p)
    plain.set_initial()
    plain.step(50)
    plan, _ = _plan(cl, p)
    cl.set_terminator(plan, cl.Terminator(p, 2, iterations=ITER))
    plan.set_initial()
    plan.step(10)
    cl.set_terminator(plan, None)
    plan.set_initial()
    plan.step(50)
    assert np.array_equal(plan.get_state(), plain.get_state())
    plain.close()
    plan.close()
