"""GPU tests of the Hermitian storage layout (N(N+1)/2 stored elements per ADO, include/helix_hip.h).

With real exponents every ADO of a Hermitian initial state stays Hermitian (checked on the oracle in
tests/test_oracle.py), so the layout must reproduce the full layout and the oracle to the plain-form bar
(1e-12 normwise relative), refuse non-Hermitian states, and be chosen only where it applies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FORM_TOL = 1e-6


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


def _rand_state(H, N, seed, hermitian=True):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((H, N, N)) + 1j * rng.standard_normal((H, N, N))
    return s + np.conj(np.transpose(s, (0, 2, 1))) if hermitian else s


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_layout_selection(hx):
    assert hx.Plan(hx.Model.spin_glass(3, 0, L=2)).layout() == "hermitian"
    assert hx.Plan(hx.Model.spin_glass(6, 1, L=1)).layout() == "hermitian"
    assert hx.Plan(hx.Model.spin_glass(7, 2, L=1)).layout() == "hermitian"      # N = 128: row tiles
    assert hx.Plan(hx.Model.legacy(N=32)).layout() == "full"                     # bit-flip coupling
    assert hx.Plan(hx.Model.spin_glass(4, 5, L=2).use_dense_hamiltonian()).layout() == "full"
    with pytest.raises(hx.HelixError):
        hx.Plan(hx.Model.legacy(N=32), layout="hermitian")


@pytest.mark.parametrize("spins,L", [(3, 4), (4, 3), (5, 2), (6, 2)])
def test_hermitian_equals_full_layout_trajectory(hx, oracle_mod, spins, L):
    """N = 8..64: 300 steps in both layouts from |1..1><1..1| and from a random Hermitian state."""
    m = hx.Model.spin_glass(spins, 3, L=L, dt=0.01)
    p = oracle_mod.spin_glass_problem(spins, 3, L=L)
    s0 = _rand_state(m.H, m.N, 5) * 1e-2
    s0[0] = p.initial()[0]
    out = {}
    for layout in ("hermitian", "full"):
        plan = hx.Plan(m, layout=layout)
        plan.set_state(s0)
        plan.step(300)
        out[layout] = plan.get_state()
        plan.close()
    assert _rel(out["hermitian"], out["full"]) <= FP64_TOL
    if m.H * m.N * m.N <= 64 * 64 * 100:
        ref = p.taylor(s0, 0.01, 300)
        assert _rel(out["hermitian"], ref) <= FP64_TOL
    h = out["hermitian"]
    assert np.max(np.abs(h - np.conj(np.transpose(h, (0, 2, 1))))) == 0.0  # unpacked state is Hermitian


@pytest.mark.parametrize("spins,L,steps", [(7, 2, 40), (8, 1, 20), (8, 2, 5)])
def test_hermitian_row_tiles_equal_full_layout(hx, spins, L, steps):
    """N = 128, 256 (the tiled kernel's 16-mode limit caps the spin glass at 8 spins): row tiles of the
    Hermitian layout (each band's rows completed in LDS from the transposed rows, far row flips gathered
    from the packed block) against the full layout."""
    m = hx.Model.spin_glass(spins, 2, L=L, dt=0.01, lam=0.3, gamma=0.7, T=1.3)
    s0 = _rand_state(m.H, m.N, 17) * 1e-2
    out = {}
    for layout in ("hermitian", "full"):
        plan = hx.Plan(m, layout=layout)
        assert plan.layout() == layout
        plan.set_state(s0)
        if layout == "hermitian":
            assert np.array_equal(plan.get_state(), s0)
        r = plan.rhs()
        plan.set_state(s0)
        plan.step(steps)
        out[layout] = (r, plan.get_state())
        plan.close()
    assert _rel(out["hermitian"][0], out["full"][0]) <= FP64_TOL
    assert _rel(out["hermitian"][1], out["full"][1]) <= FP64_TOL


@pytest.mark.parametrize("spins,L,precision,layout", [
    (3, 4, "fp64", "hermitian"), (5, 2, "fp64", "hermitian"), (6, 2, "fp64", "hermitian"), (6, 2, "fp64", "full"),
    (6, 2, "fp32", "hermitian"), (6, 2, "split_fp32", "hermitian"), (4, 3, "fp32", "hermitian"),
    (6, 2, "fp32", "full")])
def test_workgroup_shapes_agree_bitwise(hx, monkeypatch, spins, L, precision, layout):
    """The stage kernel's two workgroup shapes for whole-block tiles (512 and 1024 threads, forced with
    HELIX_HIP_TILE_THREADS; the plan's autotuner otherwise keeps the faster) compute every element with the
    same operations in the same order, so their trajectories agree bit for bit."""
    m = hx.Model.spin_glass(spins, 4, L=L, dt=0.01, lam=0.2, precision=precision)
    s0 = _rand_state(m.H, m.N, 23) * 1e-2
    out = {}
    for th in ("512", "1024"):
        monkeypatch.setenv("HELIX_HIP_TILE_THREADS", th)
        plan = hx.Plan(m, layout=layout)
        plan.set_state(s0)
        plan.step(12)
        out[th] = plan.get_state()
        plan.close()
    assert np.array_equal(out["512"], out["1024"]), _rel(out["512"], out["1024"])
    assert np.all(np.isfinite(out["512"])) and np.max(np.abs(out["512"])) > 0


@pytest.mark.parametrize("precision", ["fp32", "split_fp32"])
def test_hermitian_layout_precision_forms(hx, oracle_mod, precision):
    m = hx.Model.spin_glass(3, 0, L=4, dt=0.01, precision=precision)
    p = oracle_mod.spin_glass_problem(3, 0, L=4)
    plan = hx.Plan(m)
    assert plan.layout() == "hermitian"
    plan.set_initial()
    rho, diag = plan.step(1000, record_every=250, record_rho=True, record_diag=True)
    r = p.initial()
    for k in range(4):
        r = p.taylor(r, 0.01, 250)
        assert float(np.max(np.abs(rho[k] - r[0]))) <= FORM_TOL
        assert np.array_equal(diag[k], np.real(np.diag(rho[k])))
    assert float(np.max(np.abs(plan.tier0() - r[0]))) <= FORM_TOL


def test_non_hermitian_state_is_refused(hx):
    m = hx.Model.spin_glass(3, 0, L=2)
    s = _rand_state(m.H, m.N, 7, hermitian=False)
    plan = hx.Plan(m)
    with pytest.raises(hx.HelixError, match="not Hermitian"):
        plan.set_state(s)
    full = hx.Plan(m, layout="full")
    full.set_state(s)
    assert np.array_equal(full.get_state(), s)


def test_set_get_round_trip_exact(hx):
    m = hx.Model.spin_glass(5, 1, L=2)
    s = _rand_state(m.H, m.N, 9)
    plan = hx.Plan(m)
    plan.set_state(s)
    assert np.array_equal(plan.get_state(), s)
    assert np.array_equal(plan.tier0(), s[0])


def test_batch_layouts_agree(hx):
    """A batch of uneven instances in the Hermitian layout against the same batch in the full layout,
    and bit for bit against single Hermitian-layout plans."""
    models = [hx.Model.spin_glass_random(s, dt=0.01) for s in (1000, 1003, 1005, 1010, 1017)]
    bh = hx.Batch(models, graph=True, autotune=False)
    bf = hx.Batch(models, graph=True, autotune=False, layout="full")
    assert bh.layout() == "hermitian" and bf.layout() == "full"
    for b in (bh, bf):
        b.set_initial()
        b.step(100)
    for q, m in enumerate(models):
        assert _rel(bh.state(q), bf.state(q)) <= FP64_TOL
        plan = hx.Plan(m)
        plan.set_initial()  # |N-1><N-1|, as the batch's default
        plan.step(100)
        assert np.array_equal(plan.get_state(), bh.state(q))
        plan.close()
