"""CPU tests of the physics terminator: the oracle's restatement (oracle/closure_oracle.py) pinned against an
exact Liouville-space solve and against the deeper hierarchy it stands for, and the product's host tables
(helix_amd.closure.Terminator) against the oracle."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def CO():
    from oracle import closure_oracle
    return closure_oracle


def _rand(H, N, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(H, N, N)) + 1j * rng.normal(size=(H, N, N))


def test_fixed_point_converges_to_the_exact_stationary_solve(CO):
    t = CO.spin_glass_terminator(4, 9000, 0.28, 1.03, 0.61, 2)
    x = _rand(t.small.H, t.N, 1)
    ex = t.missing(x)
    assert np.max(np.abs(t.missing(x, 40) - ex)) <= 1e-13 * np.max(np.abs(ex))
    # each round contracts by ~|Xi| / |rate|: the error falls geometrically
    errs = [np.max(np.abs(t.missing(x, k) - ex)) for k in (1, 2, 3, 4)]
    assert all(errs[i + 1] < errs[i] / 100 for i in range(3))


def test_without_xi_one_eigenbasis_round_is_exact(CO):
    # with the Xi self term removed the stationary equation is a Sylvester equation that the eigenbasis of H
    # diagonalises: one round equals the Kronecker solve
    t = CO.spin_glass_terminator(3, 2, 0.3, 0.8, 1.4, 2)
    t.xi_self = np.zeros_like(t.xi_self)
    x = _rand(t.small.H, t.N, 2)
    ex = t.missing(x)
    assert np.max(np.abs(t.missing(x, 1) - ex)) <= 1e-13 * np.max(np.abs(ex))


def test_exact_solution_satisfies_its_equation(CO):
    t = CO.spin_glass_terminator(3, 4, 0.2, 1.2, 1.0, 1)
    x = _rand(t.small.H, t.N, 3)
    H = t.H
    for (rate, src), X in zip(t.virtual, t.exact(x)):
        D = sum(w * x[s] for s, w in src)
        res = -1j * (H @ X - X @ H) + (rate + t.xi_self) * X + D
        assert np.max(np.abs(res)) <= 1e-12 * max(1.0, np.max(np.abs(D)))


def test_terminator_is_closer_to_the_deep_hierarchy_than_plain_truncation(CO):
    # 3 spins, tiers 0..2 kept; reference: plain tiers 0..5 (462 ADOs); t = 2
    from oracle import oracle as O
    kw = dict(lam=0.25, gamma=1.0, T=1.0)
    t = CO.spin_glass_terminator(3, 0, kw["lam"], kw["gamma"], kw["T"], 2)
    deep = O.spin_glass_problem(3, 0, K=1, L=5, **kw)
    ref = deep.taylor(deep.initial(), 0.01, 200)[0]
    plain = t.small.taylor(t.small.initial(), 0.01, 200)[0]
    term = t.taylor(t.small.initial(), 0.01, 200)[0]
    e_plain, e_term = np.max(np.abs(plain - ref)), np.max(np.abs(term - ref))
    print(f"max|d rho_S| vs tiers 0..5: plain {e_plain:.2e}, terminator {e_term:.2e}")
    assert e_term < e_plain / 2


@pytest.mark.parametrize("seed", [9000, 9003])
def test_host_terminator_tables_match_the_oracle(CO, seed):
    from helix_amd import closure as cl
    p = cl.Params.draw(seed)
    ot = CO.spin_glass_terminator(4, p.seed, p.lam, p.gamma, p.T, 2)
    T = cl.Terminator(p, 2, iterations=5)
    assert T.count == len(ot.virtual) == 120
    x = _rand(T.H, T.N, seed)
    ref = ot.missing(x, 5)
    assert np.max(np.abs(T.missing(x) - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_response_targets_are_one_at_the_stationary_point():
    """Pin of the learned correction's training targets: when every tier-(L+1) ADO is exactly the terminator's
    stationary solution of the kept state (fixed point iterated to convergence), the least-squares response
    c* = -(sum conj(R~) X~ / sum |R~|^2)(r_u - i (E_a - E_b)) is 1 for every virtual ADO and eigen-element."""
    from helix_amd import closure as cl
    p = cl.Params.draw(9004)
    T = cl.Terminator(p, 2, iterations=80)
    src_small = np.unique(T.src_ado)
    local = {int(a): q for q, a in enumerate(src_small)}
    src_local = np.array([local[int(a)] for a in T.src_ado])
    seg = np.asarray(T.src_offsets[:-1], np.int64)
    num, den = 0.0, 0.0
    for k in range(3):
        x = _rand(T.H, T.N, 100 + k)
        X = T.virtual(x)
        a, b = cl.response_terms(cl.source_weights(T)[None], T.evecs[None], T._selfw[None], x[src_small][None],
                                 X[None], src_local, seg)
        num, den = num + a[0], den + b[0]
    c = -(num / den) * cl.physical_response(T)
    assert np.max(np.abs(c - 1.0)) <= 1e-9
    # a planted mistake (the response without the Xi self term in R) is caught
    a, b = cl.response_terms(cl.source_weights(T)[None], T.evecs[None], 0.0 * T._selfw[None], x[src_small][None],
                             X[None], src_local, seg)
    assert np.max(np.abs(-(a[0] / b[0]) * cl.physical_response(T) - 1.0)) > 1e-3


def test_correction_features_layout():
    from helix_amd import closure as cl
    p = cl.Params.draw(9001)
    T = cl.Terminator(p, 2)
    F = cl.correction_features(T)
    N = T.N
    assert F.shape == (T.count * N * N, cl.feature_dim(4)) and F.dtype == np.float32
    u, a, b = 7, 3, 11
    row = F[u * N * N + a * N + b]
    assert np.isclose(row[0], p.lam / 0.5) and np.isclose(row[1], p.gamma / 2) and np.isclose(row[2], p.T / 2)
    assert np.allclose(row[9:17], T.virtual_rows[u] / 3.0)
    assert np.isclose(row[17], T.rate[u].real / 10) and np.isclose(row[18], T.rate[u].imag / 10)
    assert np.allclose(row[19:22], [T.evals[a] / 4, T.evals[b] / 4, (T.evals[a] - T.evals[b]) / 4], atol=1e-6)
