"""Pins of the CPU FP64 oracle.

Every check here is independent of the oracle's own formulas: reference
known-answer tests (tests/unit/*, examples/outputEnergy.txt), published test
vectors, closed forms, and an analytic solution of the pure-dephasing limit.
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ PSD

def test_psd_k2_reference_values(oracle_mod):
    # tests/unit/psd_reference_tests.cu:23-28 (tolerance 1e-12, :7)
    poles, res = oracle_mod.psd(2)
    assert abs(poles[0] - 6.3059391442248085) <= 1e-12
    assert abs(poles[1] - 19.499618752922666) <= 1e-12
    assert abs(res[0] - 1.0328241810241552) <= 1e-12
    assert abs(res[1] - 5.9671758189758448) <= 1e-12


def test_psd_k1_closed_form(oracle_mod):
    # 2x2 tridiagonal 1/sqrt(3*5): eigenvalue 1/sqrt(15) -> xi = 2 sqrt(15); residue = K b(K+1)/2 = 2.5
    poles, res = oracle_mod.psd(1)
    assert abs(poles[0] - 2.0 * math.sqrt(15.0)) <= 1e-13
    assert abs(res[0] - 2.5) <= 1e-14


def test_psd_matches_bose_function_expansion(oracle_mod):
    # The [K-1/K] Pade expansion of 1/(1-exp(-x)) - 1/x - 1/2 ... checked the way the PSD is used:
    # x/(exp(x)-1) ~ 1 - x/2 + sum_k 2 eta_k x^2/(x^2 + xi_k^2); its error must fall quickly with K.
    xs = np.linspace(0.1, 8.0, 40)
    exact = xs / np.expm1(xs)
    errs = []
    for K in (1, 2, 3, 4, 6):
        poles, res = oracle_mod.psd(K)
        approx = 1.0 - xs / 2.0 + sum(2.0 * r * xs ** 2 / (xs ** 2 + p ** 2) for p, r in zip(poles, res))
        errs.append(np.max(np.abs(approx - exact)))
    assert errs[-1] < 1e-12
    assert all(errs[i + 1] < errs[i] / 10 for i in range(4))


# ------------------------------------------------------------ hierarchy

def test_hierarchy_known_sizes(oracle_mod):
    # tests/unit/initialize_detail_tests.cu:82-87 and C(M+L, L) for the workloads
    assert oracle_mod.hierarchy_size(3, 3) == 10
    assert oracle_mod.hierarchy_size(6, 5) == math.comb(10, 4) == 210
    assert oracle_mod.hierarchy_size(12, 6) == math.comb(17, 5) == 6188
    assert oracle_mod.hierarchy_size(16, 4) == math.comb(19, 3) == 969
    assert oracle_mod.hierarchy_size(8, 3) == math.comb(10, 2) == 45
    assert oracle_mod.hierarchy_size(8, 7) == math.comb(14, 6) == 3003


@pytest.mark.parametrize("modes,tiers", [(1, 4), (2, 4), (3, 3), (4, 5), (6, 5)])
def test_hierarchy_rows_and_edges(oracle_mod, modes, tiers):
    rows = oracle_mod.hierarchy_rows(modes, tiers)
    H = rows.shape[0]
    # brute force enumeration in lexicographic order, mode 0 most significant
    import itertools
    brute = [t for t in itertools.product(range(tiers), repeat=modes) if sum(t) < tiers]
    assert [tuple(r) for r in rows] == sorted(brute)
    edges = oracle_mod.hierarchy_edges(rows)
    index = {tuple(r): i for i, r in enumerate(rows)}
    for i, r in enumerate(rows):
        for k in range(modes):
            up = list(r); up[k] += 1
            dn = list(r); dn[k] -= 1
            assert edges[i, k] == index.get(tuple(up), H)
            assert edges[i, modes + k] == index.get(tuple(dn), H)
            if edges[i, k] != H:  # reciprocity (initialize_detail_tests.cu:105-123)
                assert edges[edges[i, k], modes + k] == i


# ----------------------------------------------------------- legacy model

def test_legacy_spin_count_and_masks(oracle_mod):
    # tests/unit/initialize_detail_tests.cu:27-36
    L = oracle_mod.lib()
    assert L.hx_legacy_spin_count(1024) == 10
    h, v = oracle_mod.legacy_operators(1024)
    row0 = set(v.cols[v.offsets[0]:v.offsets[1]].tolist())
    assert row0 == {0, 1, 2, 4, 8, 16, 32, 64, 128, 256}  # 512 mask omitted
    assert v.offsets[-1] == 1024 * 10


def test_legacy_coupling_counts(oracle_mod):
    # tests/unit/initialize_detail_tests.cu:38-56
    def count(spins, sq):
        c = np.zeros(spins * spins)
        oracle_mod.lib().hx_legacy_coupling_matrix(spins, 0.6, int(sq), oracle_mod._pd(c))
        return int(np.sum(np.abs(c) > 1e-12)), c.reshape(spins, spins)
    n, lin = count(4, False)
    assert n == 6 and lin[0, 1] == 0.6 and lin[1, 0] == 0.6 and lin[0, 2] == 0.0
    n, tri = count(9, False)
    assert n == 32 and tri[1, 7] == 0.6 and tri[7, 1] == 0.6
    n, sq = count(9, True)
    assert n == 24 and sq[1, 7] == 0.0 and sq[1, 8] == 0.6


def test_legacy_hamiltonian_diagonal_state0(oracle_mod):
    # tests/unit/initialize_detail_tests.cu:63-72: -1.5 + 0.6*0.25*2 = -1.2
    v = oracle_mod.lib().hx_legacy_hamiltonian_diagonal(0, 3, 0.6, 0, 1.0)
    assert abs(v - (-1.2)) <= 1e-12


def test_energy_helper_known_answers(oracle_mod):
    # tests/unit/energy_observable_tests.cpp:10-41: real parts only, 3.5
    L = oracle_mod.lib()
    hd = np.array([[2.0, 9.0], [0.0, 0.0], [0.0, 0.0], [4.0, 9.0]]).ravel()
    rho = np.zeros((2, 2), np.complex128)
    rho[0, 0] = 0.25 + 7j
    rho[1, 1] = 0.75 + 7j
    e = L.hx_energy_diagonal(oracle_mod._pd(hd), None, 2, oracle_mod._pd(rho.view(np.float64)))
    assert abs(e - 3.5) <= 1e-12


def test_legacy_initial_energy_closed_form(oracle_mod):
    # all spins up: sum 10 * 0.5 + 9 bonds * 0.6 * 0.25 = 6.35 (examples/outputEnergy.txt row 0: 6.3499999)
    p = oracle_mod.legacy_problem()
    assert abs(p.energy_diagonal(p.initial()) - 6.35) <= 1e-12


def test_legacy_golden_energy_trace_prefix(oracle_mod):
    # examples/outputEnergy.txt, abs tol 1e-5 (tests/integration/helix_smoke_integration.cmake:14-61)
    t, e = oracle_mod.read_energy_trace(os.path.join(GOLDEN, "legacy_outputEnergy.txt"))
    assert len(e) == 1981 and abs(e[1000] - (-1.5686448)) < 1e-9 and abs(e[198 * 10] - (-2.2040629)) < 1e-9
    p = oracle_mod.legacy_problem()
    r = p.initial()
    for step in (1, 2):
        r = p.taylor(r, 0.1, 1)
        assert abs(p.energy_diagonal(r) - e[step]) <= 1e-5
        assert abs(t[step] - 0.1 * step) < 1e-9


# ------------------------------------------------------------ RHS pins

def test_diagonal_hamiltonian_commutator_kat(oracle_mod):
    # tests/cuda/diagonal_hamiltonian_commutator_tests.cu:72-160 on the synthetic rho, n = 4 and 8
    for n in (4, 8):
        diag = np.array([-1.25 + (1 if i == 2 else i) * 0.83 for i in range(n)])
        r = np.arange(n)[:, None]
        c = np.arange(n)[None, :]
        rho = (0.25 + r * 0.37 - c * 0.11) + 1j * (-0.5 + r * 0.07 + c * 0.19)
        zero_v = oracle_mod.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.zeros(n, complex))
        p = oracle_mod.Problem(n, [oracle_mod.Bath(1.0, 1.0, 0.0, 1, zero_v)], 1, h_diag=diag)
        out = p.rhs(rho[None])[0]
        expected = -1j * (diag[:, None] * rho) + 1j * (rho * diag[None, :])
        assert np.max(np.abs(out - expected)) <= 1e-14
        assert out[0, 0] == 0 and out[1, 2] == 0  # diagonal and degenerate pair vanish


@pytest.mark.parametrize("which", ["legacy32", "tfim3"])
def test_rhs_trace_and_hermiticity(oracle_mod, which):
    # tests/numerical/sparse_drho_reference_tests.cu:53-105: tr d(rho_0) = 0 and d(rho_0) Hermitian
    if which == "legacy32":
        p = oracle_mod.legacy_problem(N=32)
    else:
        p = oracle_mod.spin_glass_problem(3, seed=0, L=3)
    rng = np.random.default_rng(1)
    rho = rng.standard_normal((p.H, p.N, p.N)) + 1j * rng.standard_normal((p.H, p.N, p.N))
    rho = rho + np.conj(np.transpose(rho, (0, 2, 1)))  # every ADO Hermitian
    d = p.rhs(rho)
    assert abs(np.trace(d[0])) <= 1e-12 * np.max(np.abs(d[0]))
    assert np.max(np.abs(d[0] - d[0].conj().T)) <= 1e-12 * np.max(np.abs(d[0]))


@pytest.mark.parametrize("which", ["legacy32", "tfim4_mixed_baths"])
def test_every_ado_stays_hermitian(oracle_mod, which):
    """The premise of the device's Hermitian storage layout: with real exponents L maps Hermitian ADOs to
    Hermitian ADOs at every tier (down_comm imaginary, down_anti real, KXi real, KXiCounter imaginary), so
    a Hermitian initial state keeps every ADO Hermitian."""
    if which == "legacy32":
        p = oracle_mod.legacy_problem(N=32)
    else:
        p = oracle_mod.spin_glass_problem(4, seed=2, L=3, bath_params=[(0.05, 0.5, 0.5), (0.4, 1.7, 1.5),
                                                                       (0.2, 1.0, 0.6), (0.3, 0.6, 1.9)])
    assert np.all(np.real(p.kxic) == 0) and np.all(np.imag(p.kxi) == 0)
    rng = np.random.default_rng(3)
    rho = rng.standard_normal((p.H, p.N, p.N)) + 1j * rng.standard_normal((p.H, p.N, p.N))
    rho = rho + np.conj(np.transpose(rho, (0, 2, 1)))
    d = p.rhs(rho)
    assert np.max(np.abs(d - np.conj(np.transpose(d, (0, 2, 1))))) <= 1e-13 * np.max(np.abs(d))
    r = p.taylor(p.initial(), 0.05, 40)
    assert np.max(np.abs(r[1:])) > 1e-3  # the hierarchy is populated
    assert np.max(np.abs(r - np.conj(np.transpose(r, (0, 2, 1))))) <= 1e-14


def test_taylor4_equals_classical_rk4(oracle_mod):
    # src/liouville.cu:502-591: for a linear L the Taylor-4 recurrence is classical RK4
    p = oracle_mod.spin_glass_problem(3, seed=0, L=2)
    r0 = p.initial()
    a = p.taylor(r0, 0.01, 5, order=4)
    b = p.rk4(r0, 0.01, 5)
    assert oracle_mod.normwise_rel(a, b) <= 1e-14


def test_unitary_limit_matches_liouville_space_taylor(oracle_mod):
    # zeta = 0 decouples every bath: rho_S must follow the degree-4 Taylor polynomial of the
    # Liouville superoperator -i(I (x) H - H^T (x) I), built independently here, and approach expm.
    from scipy.linalg import expm
    spins = 3
    J = oracle_mod.tfim_couplings(spins, 0)
    Hc = oracle_mod.tfim_hamiltonian(spins, J, 1.0)
    N = 1 << spins
    baths = [oracle_mod.Bath(1.0, 1.0, 0.0, 1, oracle_mod.sigma_z(spins, s)) for s in range(spins)]
    p = oracle_mod.Problem(N, baths, 3, h_csr=Hc)
    Hd = Hc.dense()
    I = np.eye(N)
    Lsup = -1j * (np.kron(I, Hd) - np.kron(Hd.T, I))  # column-stacking vec
    dt, steps = 0.01, 50
    P = sum(np.linalg.matrix_power(dt * Lsup, j) / math.factorial(j) for j in range(5))
    rho0 = p.initial()
    vec = rho0[0].reshape(-1, order="F")
    for _ in range(steps):
        vec = P @ vec
    expect = vec.reshape(N, N, order="F")
    got = p.taylor(rho0, dt, steps)[0]
    assert np.max(np.abs(got - expect)) <= 1e-13
    U = expm(-1j * Hd * dt * steps)
    assert np.max(np.abs(got - U @ rho0[0] @ U.conj().T)) <= 2e-7  # Taylor-4 truncation, O(dt^5) per step


def _pure_dephasing_exact(oracle_mod, e, params, rho0, t):
    """Exact sigma^z pure dephasing: rho_ij(t) = rho_ij(0) exp(-i w_ij t) prod_b exp(-(s_i-s_j)(s_i g_b - s_j g_b*))
    with g(t) = sum_k c_k (nu_k t - 1 + exp(-nu_k t)) / nu_k^2 + delta t and the standard Drude-Lorentz
    Pade coefficients c_0 = lam gam (cot(beta gam/2) - i), c_k = 4 lam gam eta_k nu_k / (beta (nu_k^2 - gam^2))."""
    N = len(e)
    out = rho0.astype(complex).copy()
    for b, (lam, gam, T, K) in enumerate(params):
        beta = 1.0 / T
        poles, res = oracle_mod.psd(K)
        nus = [gam] + [pk / beta for pk in poles]
        cs = [lam * gam * (1.0 / math.tan(beta * gam / 2) - 1j)]
        cs += [4 * lam * gam * ek * nk / (beta * (nk ** 2 - gam ** 2)) for ek, nk in zip(res, nus[1:])]
        delta = 2 * lam / (beta * gam) - lam / math.tan(beta * gam / 2) - sum(c / n for c, n in zip(cs[1:], nus[1:]))
        g = sum(c * (n * t - 1 + math.exp(-n * t)) / n ** 2 for c, n in zip(cs, nus)) + delta * t
        for i in range(N):
            for j in range(N):
                si = 1 if (i >> b) & 1 else -1
                sj = 1 if (j >> b) & 1 else -1
                out[i, j] *= np.exp(-(si - sj) * (si * g - sj * np.conj(g)))
    for i in range(N):
        for j in range(N):
            out[i, j] *= np.exp(-1j * (e[i] - e[j]) * t)
    return out


def test_pure_dephasing_two_baths_analytic(oracle_mod):
    # h = 0: H and every V_b diagonal, so the reduced dynamics has the closed form above.  The HEOM
    # truncation error must shrink geometrically with depth L towards it (1e-6 at L=8, 1e-8 at L=10).
    spins = 2
    J = np.zeros((2, 2)); J[0, 1] = 0.7
    Hc = oracle_mod.tfim_hamiltonian(spins, J, 0.0)
    params = [(0.05, 1.0, 1.0, 1), (0.08, 0.7, 1.5, 1)]
    psi = np.ones(4) / 2
    dt, steps = 0.01, 300
    exact = _pure_dephasing_exact(oracle_mod, np.real(np.diag(Hc.dense())), params, np.outer(psi, psi), dt * steps)
    errs = []
    for L in (4, 6, 8, 10):
        baths = [oracle_mod.Bath.drude_lorentz(l, g, T, K, oracle_mod.sigma_z(spins, s))
                 for s, (l, g, T, K) in enumerate(params)]
        p = oracle_mod.Problem(4, baths, L + 1, h_csr=Hc)
        rho = p.zeros(); rho[0] = np.outer(psi, psi)
        errs.append(np.max(np.abs(p.taylor(rho, dt, steps)[0] - exact)))
    assert errs[2] < 3e-6 and errs[3] < 5e-8
    assert all(errs[i + 1] < errs[i] / 10 for i in range(3))


# -------------------------------------------------------------- RNG / J

def test_splitmix64_published_vector(oracle_mod):
    # SplitMix64 reference output for seed 1234567 (S. Vigna's splitmix64.c test vector)
    import ctypes as C
    s = C.c_ulonglong(1234567)
    got = [oracle_mod.lib().hx_splitmix64(C.byref(s)) for _ in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def test_spin_glass_couplings_fixture(oracle_mod):
    # committed J matrices for seeds 0-4 (tests/golden/tfim_couplings.json)
    import json
    with open(os.path.join(GOLDEN, "tfim_couplings.json")) as f:
        fx = json.load(f)
    for key, entry in fx["couplings"].items():
        J = oracle_mod.tfim_couplings(entry["spins"], entry["seed"])
        assert np.array_equal(J, np.array(entry["J"]))


# ---------------------------------------------- TFIM Hamiltonian, <H>, normal draws
#
# hx_tfim_hamiltonian builds H for every spin-glass workload and hx_trace_h_rho gives <H>(t); both are
# pinned here against constructions that share nothing with them: Kronecker products of 2x2 Pauli
# matrices, numpy's trace of a dense product, and spectra known in closed form.

_X = np.array([[0.0, 1.0], [1.0, 0.0]])
_Z_BIT = np.diag([-1.0, 1.0])  # basis bit 1 = spin up = +1 (src/initialize_detail.h:118-121)


def _on_spin(op, s, spins):
    """op acting on spin s (basis bit s) of `spins` spins: bit s is factor spins-1-s of the Kronecker chain."""
    out = np.eye(1)
    for f in range(spins - 1, -1, -1):
        out = np.kron(out, op if f == s else np.eye(2))
    return out


def _tfim_kron(J, h):
    spins = J.shape[0]
    N = 1 << spins
    H = np.zeros((N, N))
    for i in range(spins):
        for j in range(i + 1, spins):
            H += J[i, j] * _on_spin(_Z_BIT, i, spins) @ _on_spin(_Z_BIT, j, spins)
        H += h * _on_spin(_X, i, spins)
    return H


@pytest.mark.parametrize("spins,seed,h", [(3, 0, 1.0), (4, 4, 1.0), (6, 1, 0.7)])
def test_tfim_hamiltonian_equals_kronecker_construction(oracle_mod, spins, seed, h):
    J = oracle_mod.tfim_couplings(spins, seed)
    got = oracle_mod.tfim_hamiltonian(spins, J, h).dense()
    want = _tfim_kron(J, h)
    assert np.max(np.abs(got.imag)) == 0.0
    assert np.max(np.abs(got.real - want)) <= 1e-14
    # and with couplings that are not the seeded ones (a transposed-index or sign slip shows here)
    rng = np.random.default_rng(spins)
    J2 = np.triu(rng.normal(size=(spins, spins)), 1)
    assert np.max(np.abs(oracle_mod.tfim_hamiltonian(spins, J2, h).dense().real - _tfim_kron(J2, h))) <= 1e-14


def test_tfim_hamiltonian_closed_form_spectra(oracle_mod):
    spins = 5
    # J = 0: h sum_i X_i has eigenvalues h (spins - 2k) with multiplicity C(spins, k)
    H = oracle_mod.tfim_hamiltonian(spins, np.zeros((spins, spins)), 0.8).dense().real
    ev = np.sort(np.linalg.eigvalsh(H))
    want = np.sort(np.concatenate([[0.8 * (spins - 2 * k)] * math.comb(spins, k) for k in range(spins + 1)]))
    assert np.max(np.abs(ev - want)) <= 1e-12
    # h = 0: diagonal, entry of basis state b = sum_{i<j} J_ij s_i s_j with s = 2 bit - 1
    J = oracle_mod.tfim_couplings(spins, 3)
    H = oracle_mod.tfim_hamiltonian(spins, J, 0.0).dense().real
    assert np.count_nonzero(H - np.diag(np.diag(H))) == 0
    for b in (0, 1, 6, 19, 31):
        s = np.array([2 * ((b >> q) & 1) - 1 for q in range(spins)], float)
        assert abs(H[b, b] - sum(J[i, j] * s[i] * s[j] for i in range(spins) for j in range(i + 1, spins))) <= 1e-14


def test_trace_h_rho_equals_numpy_trace(oracle_mod):
    rng = np.random.default_rng(7)
    for spins in (3, 4, 6):
        N = 1 << spins
        J = oracle_mod.tfim_couplings(spins, spins)
        H = oracle_mod.tfim_hamiltonian(spins, J, 1.0)
        rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))  # not Hermitian: Im part counts too
        p = oracle_mod.spin_glass_problem(spins, spins, L=1)
        out = np.zeros(2)
        r = np.ascontiguousarray(rho, np.complex128)
        s = H._struct()
        oracle_mod.lib().hx_trace_h_rho(oracle_mod.C.byref(s), oracle_mod._pd(r.view(np.float64)), oracle_mod._pd(out))
        want = np.trace(_tfim_kron(J, 1.0) @ rho)
        assert abs(out[0] - want.real) <= 1e-12 * N and abs(out[1] - want.imag) <= 1e-12 * N
        assert abs(p.energy(rho[None]) - want.real) <= 1e-12 * N
    # a complex Hermitian H: the imaginary parts of H enter with the right sign
    N = 8
    A = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    Hc = A + A.conj().T
    rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    out = np.zeros(2)
    Hcsr = oracle_mod.Csr.from_dense(Hc)  # kept alive: the struct points into its arrays
    s = Hcsr._struct()
    r = np.ascontiguousarray(rho, np.complex128)
    oracle_mod.lib().hx_trace_h_rho(oracle_mod.C.byref(s), oracle_mod._pd(r.view(np.float64)), oracle_mod._pd(out))
    want = np.trace(Hc @ rho)
    assert abs(out[0] - want.real) <= 1e-12 and abs(out[1] - want.imag) <= 1e-12


def test_normal_draws_have_standard_normal_moments(oracle_mod):
    # hx_normal (Box-Muller over SplitMix64): sample moments of 2e5 draws against N(0, 1), and a
    # Kolmogorov-Smirnov test; tolerances are ~5 standard errors of each estimator
    from scipy import stats
    r = oracle_mod.Rng(20240601)
    n = 200_000
    x = np.array([r.normal() for _ in range(n)])
    assert abs(x.mean()) < 5 / math.sqrt(n)
    assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert abs(stats.kurtosis(x, fisher=False) - 3.0) < 5 * math.sqrt(24.0 / n)
    assert abs(stats.skew(x)) < 5 * math.sqrt(6.0 / n)
    assert stats.kstest(x, "norm").pvalue > 1e-4
    # the couplings use consecutive draws, upper triangle only, row by row
    J = oracle_mod.tfim_couplings(4, 11)
    r2 = oracle_mod.Rng(11)
    want = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1, 4):
            want[i, j] = r2.normal()
    assert np.array_equal(J, want)
