"""CPU float64 restatement of the physics terminator of a truncated hierarchy (include/helix_hip.h,
helix_hip_plan_set_terminator).

TEST INFRASTRUCTURE ONLY: the checker for the HIP terminator kernel and the host code that builds its
tables.  Everything here is derived from the oracle's own problem description (oracle.Problem: multi-index
rows, bath constants, dense H, sigma^z couplings), not from the product's tables:

For a hierarchy kept to tiers 0..L, every tier-(L+1) ADO u of the deeper hierarchy is replaced by the
stationary solution of its own equation with its up-neighbours dropped (the reference's RHS, src/liouville.cu:
783-858, for ADO u):
    0 = -i[H, X_u] + c_Sigma(u) X_u - sum_b KXi_b [V_b, [V_b, X_u]] - sum_b KXiCounter_b {V_b, [V_b, X_u]}
        + sum_{l: u_l >= 1} ( c_psi / c_thetaC(u, l) [V_b(l), rho_{u-e_l}] + c_thetaA(u, l) {V_b(l), rho_{u-e_l}} )
with the coefficient slots of src/liouville.cu:723-744 evaluated at u.  Tier-L ADO n then gains
(-i)[V_b(k), X_{n+e_k}] for every mode k.  `exact` solves that linear system in Liouville space (Kronecker form);
`fixed_point` restates the kernel's algorithm (eigenbasis of H, a fixed number of rounds for the Xi term) and is
pinned against `exact` by the tests.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def _vdiag(bath) -> np.ndarray:
    d = bath.V.dense()
    assert np.count_nonzero(d - np.diag(np.diag(d))) == 0, "sigma^z couplings only"
    return np.real(np.diag(d))


class OracleTerminator:
    def __init__(self, small: O.Problem, big: O.Problem):
        assert small.M == big.M and small.N == big.N and big.tiers == small.tiers + 1
        self.small, self.big = small, big
        self.L = small.tiers - 1
        N, M = small.N, small.M
        self.N = N
        self.H = small.h_dense()
        v = [_vdiag(b) for b in small.baths]
        self.dv = np.stack([v[small.mode_bath[k]][:, None] - v[small.mode_bath[k]][None, :] for k in range(M)])
        self.sv = np.stack([v[small.mode_bath[k]][:, None] + v[small.mode_bath[k]][None, :] for k in range(M)])
        xi = np.zeros((N, N), np.complex128)
        for b in range(small.B):
            d = v[b][:, None] - v[b][None, :]
            s = v[b][:, None] + v[b][None, :]
            xi -= small.kxi[b] * d * d + small.kxic[b] * s * d
        self.xi_self = xi
        pos_small = {tuple(r): i for i, r in enumerate(small.rows)}
        self.virtual = []  # (rate, [(small index, weight N x N)])
        self.vpos = {}
        for r in big.rows:
            if int(r.sum()) != self.L + 1:
                continue
            rate = -sum(big.nu[k] * r[k] for k in range(M))
            src = []
            for l in range(M):
                if r[l] == 0:
                    continue
                s = r.copy()
                s[l] -= 1
                kind, b = small.mode_kind[l], small.mode_bath[l]
                if kind == 0:  # Drude mode: c_thetaC [V, .] + c_thetaA {V, .}
                    comm = -small.nu[l] * r[l] * small.kthc[b]
                    anti = -small.nu[l] * r[l] * small.ktha[b]
                else:          # Pade mode: c_psi [V, .]
                    comm = -small.nu[l] * r[l] * small.kphi[l]
                    anti = 0.0
                src.append((pos_small[tuple(s)], comm * self.dv[l] + anti * self.sv[l]))
            self.vpos[tuple(r)] = len(self.virtual)
            self.virtual.append((rate, src))
        self.up = []  # (small index n, mode k, virtual index)
        for n, r in enumerate(small.rows):
            if int(r.sum()) != self.L:
                continue
            for k in range(M):
                u = r.copy()
                u[k] += 1
                self.up.append((n, k, self.vpos[tuple(u)]))

    def _D(self, x):
        return [sum(w * x[s] for s, w in src) for _, src in self.virtual]

    def exact(self, x) -> list:
        """X_u from the Liouville-space linear system (row-major vec: vec(A X B) = (A kron B^T) vec(X))."""
        N = self.N
        I = np.eye(N)
        adH = np.kron(self.H, I) - np.kron(I, self.H.T)
        out = []
        for (rate, _), D in zip(self.virtual, self._D(x)):
            K = -1j * adH + np.diag(rate + self.xi_self.reshape(-1))
            out.append(np.linalg.solve(K, -D.reshape(-1)).reshape(N, N))
        return out

    def fixed_point(self, x, iterations: int, corr=None) -> list:
        """The kernel's algorithm: X <- solve of (-i ad_H + rate) X = -(D + Xi X) in the eigenbasis of H; with
        corr [U][N][N] every eigen-element's response is multiplied by corr[u] (the learned correction)."""
        E, Q = np.linalg.eigh(self.H)
        out = []
        for u, ((rate, _), D) in enumerate(zip(self.virtual, self._D(x))):
            den = rate - 1j * (E[:, None] - E[None, :])
            g = (1.0 if corr is None else corr[u]) / den
            X = np.zeros_like(D)
            for _ in range(iterations):
                R = D + self.xi_self * X
                X = Q @ (-(Q.conj().T @ R @ Q) * g) @ Q.conj().T
            out.append(X)
        return out

    def missing(self, x, iterations: int | None = None, corr=None) -> np.ndarray:
        X = self.exact(x) if iterations is None else self.fixed_point(x, iterations, corr)
        out = np.zeros_like(x)
        for n, k, q in self.up:
            out[n] += -1j * self.dv[k] * X[q]
        return out

    def rhs(self, x, iterations: int | None = None, corr=None) -> np.ndarray:
        return self.small.rhs(x) + self.missing(x, iterations, corr)

    def taylor(self, x, dt: float, nsteps: int, iterations: int | None = None, order: int = 4) -> np.ndarray:
        """The Taylor-m recurrence (src/liouville.cu:537-583) with the terminator in every stage."""
        x = np.array(x, np.complex128, copy=True)
        for _ in range(nsteps):
            term = x
            acc = x.copy()
            for j in range(1, order + 1):
                term = (dt / j) * self.rhs(term, iterations)
                acc = acc + term
            x = acc
        return x


def spin_glass_terminator(spins: int, seed: int, lam: float, gamma: float, T: float, keep_tier: int) -> OracleTerminator:
    small = O.spin_glass_problem(spins, seed, lam=lam, gamma=gamma, T=T, K=1, L=keep_tier)
    big = O.spin_glass_problem(spins, seed, lam=lam, gamma=gamma, T=T, K=1, L=keep_tier + 1)
    return OracleTerminator(small, big)
