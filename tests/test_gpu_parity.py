"""GPU parity of the HIP plan against the CPU FP64 oracle, through the C-ABI.

Bars: plain FP64 form <= 1e-12 normwise relative (max|a-b| / max|b|) on the RHS, on trajectories of
rho_S and on <H>(t); the legacy spin glass additionally against the reference's golden FP32 energy
trace (examples/outputEnergy.txt, abs tol 1e-5); FP32 / split-FP32 forms within max|d rho_S| <= 1e-6
of the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
FP64_TOL = 1e-12
FORM_TOL = 1e-6


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


def _rand_state(H, N, seed):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((H, N, N)) + 1j * rng.standard_normal((H, N, N))
    return s + np.conj(np.transpose(s, (0, 2, 1)))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


SPIN_CASES = [
    dict(spins=3, seed=0, L=4),                                      # 210 ADOs, N=8
    dict(spins=4, seed=3, L=2),                                      # 45 ADOs, N=16
    dict(spins=5, seed=7, L=2, bath_params=[(0.05, 0.5, 0.5), (0.2, 1.0, 1.5), (0.5, 2.0, 2.0),
                                            (0.1, 1.5, 0.7), (0.3, 0.7, 1.1)]),
    dict(spins=6, seed=1, L=2),                                      # 91 ADOs, N=64
]


@pytest.mark.parametrize("kernel,layout", [("tiled", "hermitian"), ("tiled", "full"), ("generic", "full")])
@pytest.mark.parametrize("case", SPIN_CASES, ids=lambda c: f"spins{c['spins']}_L{c['L']}")
def test_rhs_parity_spin_glass_fp64(hx, oracle_mod, case, kernel, layout, monkeypatch):
    if kernel == "generic":
        monkeypatch.setenv("HELIX_HIP_GENERIC_STAGE", "1")
    m = hx.Model.spin_glass(**case)
    bp = case.get("bath_params")
    p = oracle_mod.spin_glass_problem(case["spins"], case["seed"], L=case["L"], bath_params=bp)
    plan = hx.Plan(m, layout=layout)
    assert plan.layout() == layout
    s = _rand_state(m.H, m.N, 11)
    plan.set_state(s)
    got = plan.rhs()
    ref = p.rhs(s)
    assert _rel(got, ref) <= FP64_TOL
    assert np.array_equal(plan.get_state(), s)  # set/get round trip is exact in FP64


@pytest.mark.parametrize("kernel", ["tiled", "generic"])
@pytest.mark.parametrize("spins,L", [(7, 2), (8, 1), (9, 1), (10, 1)])
def test_rhs_parity_large_n_fp64(hx, oracle_mod, spins, L, kernel, monkeypatch):
    """N = 128..1024: row-tiled stage (far row flips from L2) and the untiled kernel."""
    if kernel == "generic":
        monkeypatch.setenv("HELIX_HIP_GENERIC_STAGE", "1")
    m = hx.Model.spin_glass(spins, seed=2, L=L, lam=0.3, gamma=0.7, T=1.3)
    p = oracle_mod.spin_glass_problem(spins, 2, L=L, lam=0.3, gamma=0.7, T=1.3)
    plan = hx.Plan(m)
    s = _rand_state(m.H, m.N, 21)
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp32", "split_fp32"])
@pytest.mark.parametrize("spins,L", [(6, 2), (7, 2), (8, 1), (9, 1), (10, 1)])
def test_rhs_parity_precision_forms(hx, oracle_mod, spins, L, precision):
    """complex64 and split-FP32 right-hand sides against the FP64 oracle, within FP32 rounding (normwise 1e-5; a
    misplaced element is an O(1) error): whole N = 64 blocks and Hermitian row bands (N = 128, 256), whose threads
    hold pairs of consecutive stored elements, and the full-layout row tiles of N = 512, 1024."""
    m = hx.Model.spin_glass(spins, seed=2, L=L, lam=0.3, gamma=0.7, T=1.3, precision=precision)
    p = oracle_mod.spin_glass_problem(spins, 2, L=L, lam=0.3, gamma=0.7, T=1.3)
    plan = hx.Plan(m)
    assert plan.layout() == ("hermitian" if spins <= 8 else "full")
    s = _rand_state(m.H, m.N, 21)
    plan.set_state(s)
    err = _rel(plan.rhs(), p.rhs(s))
    print(f"{precision} N={m.N} {m.H} ADOs: normwise rel err {err:.2e}")
    assert err <= 1e-5
    plan.close()


@pytest.mark.parametrize("precision", ["fp32", "split_fp32"])
def test_precision_form_row_band_trajectory(hx, oracle_mod, precision):
    """7 spins (N = 128, row bands), tiers 0..1, 20 Taylor-4 steps: the precision form's rho_S, every stage's
    paired Taylor sum included, against the FP64 oracle's trajectory, max|d rho_S| <= 1e-6."""
    m = hx.Model.spin_glass(7, 2, L=1, dt=0.01, lam=0.3, gamma=0.7, T=1.3, precision=precision)
    p = oracle_mod.spin_glass_problem(7, 2, L=1, lam=0.3, gamma=0.7, T=1.3)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(20)
    ref = p.taylor(p.initial(), 0.01, 20)[0]
    err = float(np.max(np.abs(plan.tier0() - ref)))
    print(f"{precision}: N=128 {m.H} ADOs 20 steps max|d rho_S| = {err:.2e}")
    assert err <= FORM_TOL
    plan.close()


@pytest.mark.parametrize("kernel", ["tiled", "generic"])
@pytest.mark.parametrize("spins", [3, 6])
def test_depth_zero_hierarchy(hx, oracle_mod, spins, kernel, monkeypatch):
    """Tiers 0..0: one ADO and no neighbours (the smallest hierarchy; only -i[H, rho] and the Xi terms act), RHS and
    200 steps against the oracle."""
    if kernel == "generic":
        monkeypatch.setenv("HELIX_HIP_GENERIC_STAGE", "1")
    m = hx.Model.spin_glass(spins, seed=6, L=0, dt=0.01)
    assert m.H == 1
    p = oracle_mod.spin_glass_problem(spins, 6, L=0)
    plan = hx.Plan(m)
    s = _rand_state(1, m.N, 41)
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL
    plan.set_initial()
    plan.step(200)
    ref = p.taylor(p.initial(), 0.01, 200)
    assert _rel(plan.get_state(), ref) <= FP64_TOL
    plan.close()


@pytest.mark.parametrize("spins", [11, 12])
def test_rhs_parity_largest_n_fp64(hx, oracle_mod, spins):
    """N = 2048 and 4096, the largest dimensions the C-ABI admits (tier 1: 23 and 25 ADOs, 1.5 and 6.7 GB of
    state): the untiled stage kernel in the full layout, on a random non-Hermitian state."""
    print(f"largest N: spins {spins}, N {1 << spins}", flush=True)
    m = hx.Model.spin_glass(spins, seed=4, L=1, lam=0.2, gamma=1.1, T=0.9)
    p = oracle_mod.spin_glass_problem(spins, 4, L=1, lam=0.2, gamma=1.1, T=0.9)
    plan = hx.Plan(m)
    assert plan.layout() == "full"  # the Hermitian layout needs the tiled kernel, which stops at N = 1024
    rng = np.random.default_rng(spins)
    s = rng.standard_normal((m.H, m.N, m.N)) + 1j * rng.standard_normal((m.H, m.N, m.N))
    plan.set_state(s)
    got = plan.rhs()
    plan.close()
    assert _rel(got, p.rhs(s)) <= FP64_TOL


def test_precision_form_n256_full_size(hx):
    """Precision form at its full size (8 spins, L=3: 969 ADOs of 256x256, dt 0.005, 1000 steps):
    complex64 and split-FP32 against the plain FP64 GPU form, whose RHS is pinned to the oracle above."""
    ref_m = hx.Model.spin_glass(8, 2, L=3, dt=0.005, precision="fp64")
    ref = hx.Plan(ref_m)
    ref.set_initial()
    ref.step(1000)
    r64 = ref.tier0()
    ref.close()
    for prec in ("fp32", "split_fp32"):
        m = hx.Model.spin_glass(8, 2, L=3, dt=0.005, precision=prec)
        plan = hx.Plan(m)
        plan.set_initial()
        plan.step(1000)
        err = float(np.max(np.abs(plan.tier0() - r64)))
        print(f"{prec}: N=256 969 ADOs 1000 steps max|d rho_S| = {err:.3e}")
        assert err <= FORM_TOL
        plan.close()


@pytest.mark.parametrize("spins,L", [(4, 3), (5, 2), (6, 2), (7, 1), (8, 1), (8, 2)])
def test_rhs_parity_dense_hamiltonian_mfma_fp64(hx, oracle_mod, spins, L):
    """Dense H on MFMA against the oracle's CSR commutator: whole blocks (N = 16, 32, 64) and 64 x 64 output
    tiles with K staged in LDS (N = 128, 256)."""
    m = hx.Model.spin_glass(spins, seed=5, L=L, lam=0.2, gamma=0.9, T=0.8).use_dense_hamiltonian()
    p = oracle_mod.spin_glass_problem(spins, 5, L=L, lam=0.2, gamma=0.9, T=0.8)
    plan = hx.Plan(m)
    assert plan.info()["dense"]
    s = _rand_state(m.H, m.N, 31)
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL


def test_dense_hamiltonian_precision_forms_n256(hx):
    """The precision workload (8 spins, N = 256, 16 modes, tiers 0..3, 969 ADOs, dt 0.005) with -i[H, rho] on the
    tiled MFMA kernel: FP64 equals the structured FP64 path to 1e-12 normwise, complex64 and split-FP32 stay within
    1e-6 of it in rho_S, over 100 steps."""
    kw = dict(spins=8, seed=2, L=3, dt=0.005)
    ref = hx.Plan(hx.Model.spin_glass(**kw))
    ref.set_initial()
    ref.step(100)
    r64 = ref.tier0()
    s64 = ref.get_state()
    ref.close()
    for prec in ("fp64", "fp32", "split_fp32"):
        m = hx.Model.spin_glass(**kw, precision=prec).use_dense_hamiltonian()
        plan = hx.Plan(m)
        assert plan.info()["dense"]
        plan.set_initial()
        plan.step(100)
        if prec == "fp64":
            assert _rel(plan.get_state(), s64) <= FP64_TOL
        else:
            err = float(np.max(np.abs(plan.tier0() - r64)))
            print(f"dense N=256 {prec}: max|d rho_S| = {err:.3e}")
            assert err <= FORM_TOL
        plan.close()


def test_dense_hamiltonian_trajectories(hx, oracle_mod):
    """4 spins, L=2 (45 ADOs of 16x16), 500 steps: FP64 dense path vs oracle, FP32/split dense vs oracle."""
    p = oracle_mod.spin_glass_problem(4, 4, L=2)
    ref = p.taylor(p.initial(), 0.01, 500)
    for prec, tol in (("fp64", FP64_TOL), ("fp32", FORM_TOL), ("split_fp32", FORM_TOL)):
        m = hx.Model.spin_glass(4, 4, L=2, dt=0.01, precision=prec).use_dense_hamiltonian()
        plan = hx.Plan(m)
        plan.set_initial()
        plan.step(500)
        if prec == "fp64":
            assert _rel(plan.get_state(), ref) <= tol
        else:
            err = float(np.max(np.abs(plan.tier0() - ref[0])))
            print(f"dense {prec}: max|d rho_S| = {err:.3e}")
            assert err <= tol


@pytest.mark.parametrize("N", [8, 32, 1024])
def test_rhs_parity_legacy_fp64(hx, oracle_mod, N):
    m = hx.Model.legacy(N=N)
    p = oracle_mod.legacy_problem(N=N)
    plan = hx.Plan(m)
    assert plan.info()["flip"]
    s = _rand_state(m.H, m.N, 5)
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL


def test_cfg0_trajectory_fp64(hx, oracle_mod):
    """3 spins, L=4 (210 ADOs of 8x8), dt 0.01, 1000 steps: rho_S(t) and <H>(t) every 100 steps."""
    m = hx.Model.spin_glass(3, 0, L=4, dt=0.01)
    p = oracle_mod.spin_glass_problem(3, 0, L=4)
    plan = hx.Plan(m)
    plan.set_initial()
    rho_tr, _ = plan.step(1000, record_every=100, record_rho=True)
    r = p.initial()
    for k in range(10):
        r = p.taylor(r, 0.01, 100)
        assert _rel(rho_tr[k], r[0]) <= FP64_TOL
        assert abs(m.energy(rho_tr[k]) - p.energy(r)) <= FP64_TOL * max(1.0, abs(p.energy(r)))
    assert abs(np.trace(rho_tr[-1]) - 1.0) <= 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_taylor_orders(hx, oracle_mod, order):
    m = hx.Model.spin_glass(3, 0, L=3, dt=0.02, order=order)
    p = oracle_mod.spin_glass_problem(3, 0, L=3)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(25)
    ref = p.taylor(p.initial(), 0.02, 25, order=order)
    assert _rel(plan.get_state(), ref) <= FP64_TOL


def test_legacy_three_steps_fp64(hx, oracle_mod):
    m = hx.Model.legacy(N=1024)
    p = oracle_mod.legacy_problem()
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(3)
    ref = p.taylor(p.initial(), 0.1, 3)
    assert _rel(plan.get_state(), ref) <= FP64_TOL


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-5), ("fp32", 1e-5)])
def test_legacy_golden_energy_trace(hx, oracle_mod, precision, tol):
    """All 1981 rows of examples/outputEnergy.txt (row r = energy before step r, abs tol 1e-5)."""
    t, e = oracle_mod.read_energy_trace(os.path.join(GOLDEN, "legacy_outputEnergy.txt"))
    m = hx.Model.legacy(N=1024, precision=precision)
    plan = hx.Plan(m)
    plan.set_initial()
    _, diag = plan.step(len(e) - 1, record_every=1, record_diag=True)
    hd = m.tables()["h_diag"]
    energies = np.concatenate([[m.energy_diagonal(m.initial_rho())], diag @ hd])
    assert np.max(np.abs(energies - e)) <= tol


def test_cfg1_full_size_rhs_fp64(hx, oracle_mod):
    """6 spins, L=5: 6188 ADOs of 64x64 -- the benchmark workload, one RHS at full size."""
    m = hx.Model.spin_glass(6, 1, L=5, dt=0.005)
    p = oracle_mod.spin_glass_problem(6, 1, L=5)
    plan = hx.Plan(m)
    s = _rand_state(m.H, m.N, 2)
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL


def test_cfg1_full_size_invariants(hx):
    """Size-independent properties at full size: unit trace and Hermitian rho_S, tier-0 <H> finite."""
    m = hx.Model.spin_glass(6, 1, L=5, dt=0.005)
    plan = hx.Plan(m)
    plan.set_initial()
    rho, _ = plan.step(200, record_every=50, record_rho=True)
    for r in rho:
        assert abs(np.trace(r) - 1.0) <= 1e-12
        assert np.max(np.abs(r - r.conj().T)) <= 1e-12
    assert np.isfinite(m.energy(rho[-1]))


@pytest.mark.parametrize("precision", ["fp32", "split_fp32"])
def test_precision_forms_cfg0(hx, oracle_mod, precision):
    m = hx.Model.spin_glass(3, 0, L=4, dt=0.01, precision=precision)
    p = oracle_mod.spin_glass_problem(3, 0, L=4)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(1000)
    ref = p.taylor(p.initial(), 0.01, 1000)[0]
    err = float(np.max(np.abs(plan.tier0() - ref)))
    print(f"{precision}: max|d rho_S| = {err:.3e}")
    assert err <= FORM_TOL


def test_cpp_example_runs(hx):
    exe = os.path.join(ROOT, "helix_amd", "helix_example")
    out = subprocess.run([exe, "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "reduced_density_shape=(1,1024,1024)" in out.stdout
    tr = float(out.stdout.strip().split("trace=")[1])
    assert abs(tr - 1.0) <= 1e-5


def test_reference_example_compiled_unchanged_runs(hx):
    """The reference's examples/cpp/legacy_spin_glass.cpp, compiled unchanged against include/helix by the build
    (which needs the reference's file next to the tree), runs on the GPU and prints the reference's line."""
    exe = os.path.join(ROOT, "helix_amd", "helix_reference_example")
    if not os.path.exists(exe):
        pytest.skip("helix_reference_example not built (the build had no reference tree next to it)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "helix_cpp_example: steps=2 time=0.2 reduced_density_shape=(1,1024,1024)"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_chunks_match_eager_steps(hx, precision):
    """helix_hip_step replays 16-step hipGraph chunks; helix_hip_time_steps launches every stage eagerly.
    Same kernels, same arguments: the states must agree bit for bit, also across record points that split
    chunks (record_every = 20) and over several replays of one graph."""
    m = hx.Model.spin_glass(4, 2, L=3, dt=0.01, precision=precision)
    a, b = hx.Plan(m), hx.Plan(m)
    a.set_initial()
    b.set_initial()
    rho, _ = a.step(100, record_every=20, record_rho=True)
    b.time_steps(100)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(rho[-1], b.tier0())
    a.step(48)
    b.time_steps(48)
    assert np.array_equal(a.get_state(), b.get_state())
