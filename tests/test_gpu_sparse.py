"""GPU parity of the generic sparse path (System::from_sparse with operators that are neither diagonal
nor bit-flip structured): CSR H and CSR couplings, including complex Hermitian ones, against the CPU
FP64 oracle, which forms A X and X A from the CSR arrays directly (no Hermiticity shortcut), at the
plain-form bar of 1e-12 normwise relative; the C++ API runs the same systems through HEOMSolver::run."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "helix_amd", "helix_sparse_system")
FP64_TOL = 1e-12


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


def _hermitian(N, density, seed, complex_):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N)) + (1j * rng.standard_normal((N, N)) if complex_ else 0)
    a = np.where(rng.random((N, N)) < density, a, 0)
    return (a + a.conj().T) / 2


def _system(N, seed):
    H = _hermitian(N, 0.35, seed, True)
    V = [_hermitian(N, 0.3, seed + 1, False), _hermitian(N, 0.3, seed + 2, True)]
    baths = [(2.0, 1.0, 0.2), (1.0, 0.7, 0.5)]  # (beta, gamma, zeta)
    return H, V, baths


def _oracle(oracle_mod, N, H, V, baths, tiers, K=1):
    O = oracle_mod
    return O.Problem(N, [O.Bath(b, g, z, K, O.Csr.from_dense(v)) for (b, g, z), v in zip(baths, V)], tiers,
                     h_csr=O.Csr.from_dense(H))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("N", [8, 16])
def test_rhs_general_sparse_fp64(hx, oracle_mod, N):
    H, V, baths = _system(N, 3)
    m = hx.Model.sparse(H, V, baths, K=1, max_tier=2)
    p = _oracle(oracle_mod, N, H, V, baths, tiers=3)
    plan = hx.Plan(m)
    assert plan.layout() == "full" and plan.info()["launches_per_step"] == 8
    rng = np.random.default_rng(5)
    s = rng.standard_normal((m.H, N, N)) + 1j * rng.standard_normal((m.H, N, N))  # not Hermitian: full layout
    plan.set_state(s)
    assert _rel(plan.rhs(), p.rhs(s)) <= FP64_TOL


def test_trajectory_general_sparse(hx, oracle_mod):
    N = 8
    H, V, baths = _system(N, 7)
    p = _oracle(oracle_mod, N, H, V, baths, tiers=3)
    r0 = p.initial()
    ref = p.taylor(r0, 0.02, 150)
    m = hx.Model.sparse(H, V, baths, K=1, max_tier=2, dt=0.02)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(150)
    assert _rel(plan.get_state(), ref) <= FP64_TOL
    m32 = hx.Model.sparse(H, V, baths, K=1, max_tier=2, dt=0.02, precision="fp32")
    p32 = hx.Plan(m32)
    p32.set_initial()
    p32.step(150)
    assert float(np.max(np.abs(p32.tier0() - ref[0]))) <= 1e-6


def _write_system(path, N, H, V, baths, tiers, steps, dt, precision):
    def op(a):
        rows = [np.nonzero(a[i])[0] for i in range(N)]
        off = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
        cols = np.concatenate(rows) if off[-1] else np.zeros(0, int)
        vals = np.concatenate([a[i, r] for i, r in enumerate(rows)]) if off[-1] else np.zeros(0, complex)
        return (f"{off[-1]}\n" + " ".join(map(str, off)) + "\n" + " ".join(map(str, cols)) + "\n" +
                " ".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in vals) + "\n")
    with open(path, "w") as f:
        f.write(f"{N} {len(V)} {tiers} {steps} {dt!r} {precision}\n")
        f.write(op(H))
        for (b, g, z), v in zip(baths, V):
            f.write(op(v))
            f.write(f"{b!r} {g!r} {z!r} 1\n")


def test_cpp_api_runs_general_sparse_system(hx, oracle_mod, tmp_path):
    N = 8
    H, V, baths = _system(N, 11)
    path = tmp_path / "system.txt"
    _write_system(path, N, H, V, baths, tiers=3, steps=60, dt=0.05, precision=0)
    r = subprocess.run([EXE, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rho = np.zeros((N, N), complex)
    for ln in r.stdout.splitlines():
        _, i, j, re, im = ln.split()
        rho[int(i), int(j)] = float(re) + 1j * float(im)
    p = _oracle(oracle_mod, N, H, V, baths, tiers=3)
    ref = p.taylor(p.initial(), 0.05, 60)[0]
    assert _rel(rho, ref) <= FP64_TOL


def test_non_hermitian_operator_is_refused(hx):
    N = 8
    H, V, baths = _system(N, 13)
    V[0][0, 1] += 0.5  # breaks Hermiticity
    m = hx.Model.sparse(H, V, baths, K=1, max_tier=1)
    with pytest.raises(hx.HelixError, match="not Hermitian"):
        hx.Plan(m)
