"""GPU tests of the chosen forms at the sizes the benchmark measures.

- complex64 and split-FP32 on the benchmarked 6-spin workload (N = 64, 12 modes, tiers 0..5, 6188 ADOs,
  dt 0.005): max|delta rho_S| against the FP64 run from the same initial state, <= 1e-6 after 200 steps.
- the 512-instance batched workload (seeds 1000..1511: spins 3-6, tiers 2-5, random bath parameters,
  489,480 ADOs, 75 GiB FP64 in one device arena) in FP64 and complex64: bitwise equal to single plans for
  instances spread over the arena (the last ones lie beyond 4 GiB offsets), the smallest instances against
  the CPU FP64 oracle, unit trace and Hermitian rho_S for every instance, and complex64 within 1e-6 of FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BENCH = dict(spins=6, seed=1, lam=0.1, gamma=1.0, T=1.0, K=1, L=5, dt=0.005)
STEPS = 200
BATCH_STEPS = 12
SEEDS = list(range(1000, 1512))


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


@pytest.fixture(scope="module")
def fp64_bench_rho(hx):
    m = hx.Model.spin_glass(**BENCH, precision="fp64")
    p = hx.Plan(m)
    p.set_initial()
    p.step(STEPS)
    r = p.tier0()
    p.close()
    return r


@pytest.mark.parametrize("precision", ["fp32", "split_fp32"])
def test_precision_forms_at_benchmarked_size(hx, fp64_bench_rho, precision):
    print(f"forms: 6 spins, 6188 ADOs, {precision}, {STEPS} steps", flush=True)
    m = hx.Model.spin_glass(**BENCH, precision=precision)
    p = hx.Plan(m)
    p.set_initial()
    p.step(STEPS)
    err = float(np.max(np.abs(p.tier0() - fp64_bench_rho)))
    p.close()
    print(f"{precision}: max|d rho_S| = {err:.3e} after {STEPS} steps")
    assert err <= 1e-6


def _batch(hx, precision):
    models = [hx.Model.spin_glass_random(s, dt=0.01, precision=precision) for s in SEEDS]
    b = hx.Batch(models, graph=True, autotune=True)
    b.set_initial()
    b.step(BATCH_STEPS)
    return models, b


@pytest.fixture(scope="module")
def batch64(hx):
    print("batch: 512 instances FP64", flush=True)
    models, b = _batch(hx, "fp64")
    yield models, b
    b.close()


def _check_single(hx, b, models, q, precision):
    m = hx.Model.spin_glass_random(SEEDS[q], dt=0.01, precision=precision)
    p = hx.Plan(m)
    p.set_initial()
    p.step(BATCH_STEPS)
    assert np.array_equal(b.state(q), p.get_state()), f"instance {q} (seed {SEEDS[q]}, {models[q].H} ADOs)"
    assert np.array_equal(b.tier0(q), p.tier0())
    p.close()


def test_batch512_fp64_equals_single_plans(hx, batch64):
    models, b = batch64
    info = b.info()
    assert info["total_ado"] == sum(m.H for m in models) and len(models) == 512
    assert info["device_bytes"] > 4 << 30
    big = max(range(512), key=lambda q: models[q].H)
    for q in sorted({0, 1, 2, 255, big, 508, 509, 510, 511}):
        _check_single(hx, b, models, q, "fp64")


def test_batch512_fp64_smallest_against_oracle(hx, oracle_mod, batch64):
    models, b = batch64
    small = sorted(range(512), key=lambda q: (models[q].H * models[q].N ** 2, q))[:3]
    for q in small:
        pr = oracle_mod.batched_instance_params(q)
        assert pr["seed"] == SEEDS[q] and (1 << pr["spins"]) == models[q].N
        o = oracle_mod.spin_glass_problem(pr["spins"], pr["seed"], lam=pr["lam"], gamma=pr["gamma"], T=pr["T"],
                                          L=pr["L"])
        assert o.H == models[q].H
        ref = o.taylor(o.initial(), 0.01, BATCH_STEPS)
        got = b.state(q)
        assert oracle_mod.normwise_rel(got, ref) <= 1e-12, f"instance {q}"


def test_batch512_every_rho_s_unit_trace_hermitian(hx, batch64):
    models, b = batch64
    for q in range(512):
        r = b.tier0(q)
        assert abs(np.trace(r) - 1.0) <= 1e-12, q
        assert np.max(np.abs(r - r.conj().T)) <= 1e-13, q


def test_batch512_complex64_against_fp64(hx, batch64):
    models, b64 = batch64
    print("batch: 512 instances complex64", flush=True)
    m32, b32 = _batch(hx, "fp32")
    worst = max(float(np.max(np.abs(b32.tier0(q) - b64.tier0(q)))) for q in range(512))
    print(f"batched complex64, 512 instances: worst max|d rho_S| = {worst:.3e}")
    assert worst <= 1e-6
    for q in (0, 511):
        _check_single(hx, b32, m32, q, "fp32")
    b32.close()
