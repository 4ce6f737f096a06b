"""CPU tests of the native host side: the C-ABI libraries load and export every declared symbol, and
the C++17 set-up (Pade, hierarchy ranking, coefficient slots, operators, seeded couplings) equals the
oracle's independent restatement and the reference's known answers.  No GPU call is made here."""
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    from helix_amd import capi
    if not (os.path.exists(capi.HIP_LIB) and os.path.exists(capi.HOST_LIB)):
        from helix_amd import build
        build.build()
    return helix_amd


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(helix_(?:hip|model|setup|io)_\w+)\s*\(", text)))


def test_libraries_export_every_declared_symbol(hx):
    from helix_amd import capi
    hip, host = capi.hip(), capi.host()
    for name in _declared("helix_hip.h"):
        assert hasattr(hip, name), name
    for name in _declared("helix_setup.h"):
        assert hasattr(host, name), name
    for name in _declared("helix_mlp.h"):
        assert hasattr(hip, name), name
    assert set(_declared("helix_hip.h")) == set(capi.HIP_SYMBOLS)
    assert set(_declared("helix_setup.h")) == set(capi.SETUP_SYMBOLS)
    for name in _declared("helix_io.h"):
        assert hasattr(host, name), name
    assert set(_declared("helix_io.h")) == set(capi.IO_SYMBOLS)
    assert len(_declared("helix_mlp.h")) == 9


def test_pade_reference_values(hx):
    # tests/unit/psd_reference_tests.cu:23-28 at 1e-12; K=1 closed form 2 sqrt(15), 2.5
    p, r = hx.capi.pade(2)
    assert abs(p[0] - 6.3059391442248085) <= 1e-12 and abs(p[1] - 19.499618752922666) <= 1e-12
    assert abs(r[0] - 1.0328241810241552) <= 1e-12 and abs(r[1] - 5.9671758189758448) <= 1e-12
    p, r = hx.capi.pade(1)
    assert abs(p[0] - 2 * np.sqrt(15.0)) <= 1e-13 and abs(r[0] - 2.5) <= 1e-14


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5])
def test_pade_matches_oracle(hx, oracle_mod, K):
    p, r = hx.capi.pade(K)
    po, ro = oracle_mod.psd(K)
    assert np.max(np.abs(p - po) / po) <= 1e-12
    assert np.max(np.abs(r - ro) / np.abs(ro)) <= 1e-10


@pytest.mark.parametrize("spins,L", [(3, 4), (4, 2), (3, 2), (5, 3)])
def test_hierarchy_and_coefficients_match_oracle(hx, oracle_mod, spins, L):
    m = hx.Model.spin_glass(spins, seed=spins, L=L, lam=0.2, gamma=1.3, T=0.8)
    p = oracle_mod.spin_glass_problem(spins, seed=spins, L=L, lam=0.2, gamma=1.3, T=0.8)
    t = m.tables()
    assert m.H == p.H and m.M == p.M
    assert np.array_equal(m.rows(), p.rows)
    assert np.array_equal(t["up"], p.edges[:, :p.M]) and np.array_equal(t["down"], p.edges[:, p.M:])
    assert np.array_equal(t["tier"], p.rows.sum(axis=1))
    # coefficient slots of src/liouville.cu:723-744 built from the oracle's bath constants
    for n in range(p.H):
        row = p.rows[n]
        for k in range(p.M):
            b, kind = p.mode_bath[k], p.mode_kind[k]
            c = p.coefs[b]
            if kind > 0:
                exp_c, exp_a = p.kphi[k] * (-p.nu[k] * row[k]), 0j
            else:
                exp_c, exp_a = c.kthc * (-p.nu[k] * row[k]), c.ktha * (-p.nu[k] * row[k])
            assert abs(t["down_comm"][n, k] - exp_c) <= 1e-14 * max(1.0, abs(exp_c))
            assert abs(t["down_anti"][n, k] - exp_a) <= 1e-14 * max(1.0, abs(exp_a))
        assert abs(t["self_coef"][n] + np.dot(p.nu, row)) <= 1e-13
    assert np.allclose(t["xi"], p.kxi, rtol=1e-12, atol=1e-300)
    assert np.allclose(t["xi_counter"], p.kxic, rtol=1e-14)


def test_workload_sizes(hx):
    for spins, L, H in [(3, 4, 210), (6, 5, 6188), (8, 3, 969), (4, 2, 45)]:
        assert hx.capi.host().helix_setup_hierarchy_size(2 * spins, L + 1) == H


def test_spin_glass_hamiltonian_matches_oracle(hx, oracle_mod):
    for spins in (3, 4):
        m = hx.Model.spin_glass(spins, seed=0)
        t = m.tables()
        Hd = np.diag(t["h_diag"]).astype(complex)
        for mask, c in zip(t["h_masks"], t["h_coefs"]):
            for i in range(m.N):
                Hd[i, i ^ mask] += c
        ref = oracle_mod.tfim_hamiltonian(spins, oracle_mod.tfim_couplings(spins, 0), 1.0).dense()
        assert np.array_equal(Hd, ref)


def test_legacy_operators_match_oracle(hx, oracle_mod):
    for N in (32, 1024):
        m = hx.Model.legacy(N=N)
        t = m.tables()
        hd, v = oracle_mod.legacy_operators(N)
        assert np.array_equal(t["h_diag"], hd)
        p = m.problem_ptr().contents
        cp = p.couplings[0]
        vd = np.ctypeslib.as_array(cp.diag, shape=(N,))
        masks = np.ctypeslib.as_array(cp.flip_masks, shape=(cp.nflips,))
        coefs = np.ctypeslib.as_array(cp.flip_coefs, shape=(cp.nflips,))
        dense = np.diag(vd).astype(complex)
        for mk, c in zip(masks, coefs):
            for i in range(N):
                dense[i, i ^ mk] += c
        ref = v.dense()
        assert np.array_equal(dense, ref)


def test_couplings_fixture(hx):
    with open(os.path.join(GOLDEN, "tfim_couplings.json")) as f:
        fx = json.load(f)
    for entry in fx["couplings"].values():
        J = hx.capi.spin_glass_couplings(entry["spins"], entry["seed"])
        assert np.array_equal(J, np.array(entry["J"]))


def test_energy_helpers(hx, oracle_mod):
    m = hx.Model.legacy(N=1024)
    assert abs(m.energy_diagonal(m.initial_rho()) - 6.35) <= 1e-12
    s = hx.Model.spin_glass(3, seed=0)
    rng = np.random.default_rng(3)
    r = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    p = oracle_mod.spin_glass_problem(3, seed=0)
    assert abs(s.energy(r) - p.energy(r[None])) <= 1e-12
    assert abs(s.energy_diagonal(r) - p.energy_diagonal(r[None])) <= 1e-12


def test_invalid_problem_rejected_without_gpu(hx):
    # set-up errors surface as contract errors, not crashes
    with pytest.raises(hx.HelixError):
        hx.Model.spin_glass(2, seed=0)
