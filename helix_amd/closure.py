"""Learned closure of a truncated hierarchy.

A hierarchy kept to tiers 0..L loses, for every tier-L ADO n, the up-neighbour term
    c_n = -i sum_k [V_{b(k)}, rho_{n+e_k}]
(the reference's plain truncation reads a zero block there, src/liouville.cu:818-823).  For sigma^z
couplings a Markovian terminator replaces it by an elementwise rate, c_n[ij] ~ kappa_n[ij] rho_n[ij];
here kappa is learned: a 3-layer perceptron (include/helix_mlp.h: BF16 weights, FP32 accumulation,
trained on the GPU by its own HIP kernels) maps the spin configuration of the element (n_b(i), n_b(j)),
the ADO's multi-index and the model parameters (lambda, gamma, T, J_ij) to (Re kappa, Im kappa).
Training targets come from the plain FP64 solver at a deeper truncation (the batched GPU form), where
the tier-(L+1) ADOs exist: per parameter set, ADO and element, the least-squares rate over the recorded
snapshots, kappa* = sum_t conj(rho_t) c_t / sum_t |rho_t|^2.  The plan evaluates the network once
(helix_hip_plan_set_closure) and its stage kernels apply kappa in every Taylor stage; a calibration
window against a deeper plain run decides whether the closure is trusted or the solver falls back to
plain truncation.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import HelixError, Model, Plan, _check, hip

_fp = C.POINTER(C.c_float)


def _bind():
    L = hip()
    if getattr(L, "_mlp_bound", False):
        return L
    vp = C.c_void_p
    L.helix_hip_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, C.POINTER(vp)]
    L.helix_hip_mlp_param_count.argtypes = [vp]
    L.helix_hip_mlp_param_count.restype = C.c_size_t
    L.helix_hip_mlp_set_params.argtypes = [vp, _fp]
    L.helix_hip_mlp_get_params.argtypes = [vp, _fp]
    L.helix_hip_mlp_forward.argtypes = [vp, _fp, C.c_int, _fp]
    L.helix_hip_mlp_gradient.argtypes = [vp, _fp, _fp, C.c_int, _fp, _fp]
    L.helix_hip_mlp_train.argtypes = [vp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, _fp]
    L.helix_hip_mlp_destroy.argtypes = [vp]
    L.helix_hip_mlp_destroy.restype = None
    L.helix_hip_mlp_last_error.restype = C.c_char_p
    L.helix_hip_plan_set_closure.argtypes = [vp, vp, C.POINTER(C.c_int32), C.c_int, _fp, C.c_int]
    L.helix_hip_plan_set_closure_rates.argtypes = [vp, C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_double)]
    L._mlp_bound = True
    return L


def _f(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def _mcheck(rc, what):
    if rc != 0:
        raise HelixError(f"{what} failed (rc={rc}): {_bind().helix_hip_mlp_last_error().decode()}")


class MLP:
    """The closure network on the GPU (helix_hip_mlp)."""

    def __init__(self, din: int, h1: int = 256, h2: int = 256, dout: int = 1, seed: int = 7, device: int = 0):
        L = _bind()
        self.dims = (din, h1, h2, dout)
        self._m = C.c_void_p()
        _mcheck(L.helix_hip_mlp_create(din, h1, h2, dout, seed, device, C.byref(self._m)), "helix_hip_mlp_create")
        self.nparam = int(L.helix_hip_mlp_param_count(self._m))

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        _mcheck(_bind().helix_hip_mlp_get_params(self._m, _f(p)), "get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        _mcheck(_bind().helix_hip_mlp_set_params(self._m, _f(p)), "set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.dims[3]), np.float32)
        _mcheck(_bind().helix_hip_mlp_forward(self._m, _f(x), x.shape[0], _f(y)), "forward")
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_float()
        _mcheck(_bind().helix_hip_mlp_gradient(self._m, _f(x), _f(t), x.shape[0], _f(g), C.byref(loss)), "gradient")
        return float(loss.value), g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 1):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(epochs, np.float32)
        _mcheck(_bind().helix_hip_mlp_train(self._m, _f(x), _f(t), x.shape[0], epochs, batch, lr, seed, _f(losses)),
                "train")
        return losses

    @property
    def handle(self):
        return self._m

    def close(self):
        if self._m:
            _bind().helix_hip_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workload

@dataclass
class Params:
    """One 4-spin spin-glass parameter set (one sigma^z bath per spin with common lambda, gamma, T)."""
    seed: int
    spins: int
    lam: float
    gamma: float
    T: float

    @staticmethod
    def draw(seed: int, spins: int = 4) -> "Params":
        """lambda ~ U(0.05, 0.5), gamma ~ U(0.5, 2), T ~ U(0.5, 2) from SplitMix64(seed), in that order."""
        st = [seed & 0xFFFFFFFFFFFFFFFF]

        def u():
            st[0] = (st[0] + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            z = st[0]
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
            z ^= z >> 31
            return (z >> 11) * (1.0 / 9007199254740992.0)
        return Params(seed, spins, 0.05 + 0.45 * u(), 0.5 + 1.5 * u(), 0.5 + 1.5 * u())

    def model(self, max_tier: int, precision: str = "fp64", dt: float = 0.01) -> Model:
        return Model.spin_glass(self.spins, self.seed, lam=self.lam, gamma=self.gamma, T=self.T, K=1, L=max_tier,
                                dt=dt, precision=precision)

    def extra(self, rows: np.ndarray, max_tier: int) -> np.ndarray:
        """Per-ADO extra features: multi-index / L, lambda, gamma, T, J_ij (i < j)."""
        J = capi.spin_glass_couplings(self.spins, self.seed)
        iu = np.triu_indices(self.spins, 1)
        glob = np.concatenate([[self.lam, self.gamma, self.T], J[iu]])
        return np.concatenate([rows.astype(np.float64) / max_tier, np.tile(glob, (rows.shape[0], 1))],
                              axis=1).astype(np.float32)


def hamiltonian_dense(model: Model) -> np.ndarray:
    """H of a spin-structured model as a dense N x N array (diagonal plus bit flips)."""
    t = model.tables()
    N = model.N
    h = np.diag(t["h_diag"]).astype(np.complex128)
    for m, c in zip(t["h_masks"], t["h_coefs"]):
        for i in range(N):
            h[i, i ^ m] += c
    return h


class Terminator:
    """The physics terminator of a hierarchy kept to tiers 0..L (include/helix_hip.h, helix_hip_plan_set_terminator):
    the tier-(L+1) ADOs of the deeper hierarchy become virtual ADOs solved in every Taylor stage from the stationary
    condition of their own equation.  Built from the tables of the hierarchy kept to tiers 0..L+1 (its tier-(L+1)
    self rates and down coefficients) and the eigenbasis of H; `missing(x)` evaluates the same terms on the host."""

    def __init__(self, params: Params, keep_tier: int, iterations: int = 3, dt: float = 0.01):
        small = params.model(keep_tier, "fp64", dt)
        big = params.model(keep_tier + 1, "fp64", dt)
        self.N, self.H, self.M = small.N, small.H, small.M
        rs, rb = small.rows(), big.rows()
        tb = big.tables()
        pos = {tuple(r): i for i, r in enumerate(rs)}
        us = np.nonzero(rb.sum(1) == keep_tier + 1)[0]
        upos = {int(u): q for q, u in enumerate(us)}
        self.rate = np.ascontiguousarray(tb["self_coef"][us], np.complex128)
        off, sado, smode, scomm, santi = [0], [], [], [], []
        for u in us:
            for l in range(big.M):
                d = tb["down"][u, l]
                if d < big.H:
                    sado.append(pos[tuple(rb[d])])
                    smode.append(l)
                    scomm.append(tb["down_comm"][u, l])
                    santi.append(tb["down_anti"][u, l])
            off.append(len(sado))
        self.src_offsets = np.array(off, np.int32)
        self.src_ado, self.src_mode = np.array(sado, np.int32), np.array(smode, np.int32)
        self.src_comm, self.src_anti = np.array(scomm, np.complex128), np.array(santi, np.complex128)
        big_of = {tuple(r): i for i, r in enumerate(rb)}
        uado, umode, uvirt = [], [], []
        for n in np.nonzero(rs.sum(1) == keep_tier)[0]:
            bn = big_of[tuple(rs[n])]
            for k in range(big.M):
                u = tb["up"][bn, k]
                if u < big.H:
                    uado.append(n)
                    umode.append(k)
                    uvirt.append(upos[int(u)])
        self.up_ado, self.up_mode = np.array(uado, np.int32), np.array(umode, np.int32)
        self.up_virtual = np.array(uvirt, np.int32)
        self.evals, self.evecs = np.linalg.eigh(hamiltonian_dense(small))
        self.evecs = np.ascontiguousarray(self.evecs, np.complex128)
        self.iterations = iterations
        ts = small.tables()
        self.mode_bath = ts["mode_bath"]
        spins = int(round(math.log2(self.N)))
        sg = _sz_signs(spins)
        self._dv = np.stack([sg[b][:, None] - sg[b][None, :] for b in self.mode_bath])
        self._sv = np.stack([sg[b][:, None] + sg[b][None, :] for b in self.mode_bath])
        B = len(ts["xi"])
        dd = np.stack([sg[b][:, None] - sg[b][None, :] for b in range(B)])
        self._selfw = -np.einsum("b,bij->ij", ts["xi"], dd * dd)
        self.count = len(us)
        small.close()
        big.close()

    def struct(self):
        """helix_hip_terminator pointing into this object's arrays (keep the object alive while it is used)."""
        self._r = self.rate.view(np.float64)
        self._c = np.ascontiguousarray(self.src_comm).view(np.float64)
        self._a = np.ascontiguousarray(self.src_anti).view(np.float64)
        self._q = self.evecs.view(np.float64)
        self._e = np.ascontiguousarray(self.evals, np.float64)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        return capi.Terminator(self.count, dp(self._r), ip(self.src_offsets), ip(self.src_ado), ip(self.src_mode),
                               dp(self._c), dp(self._a), len(self.up_ado), ip(self.up_ado), ip(self.up_mode),
                               ip(self.up_virtual), dp(self._q), dp(self._e), self.iterations)

    def virtual(self, x: np.ndarray) -> np.ndarray:
        """X_u of every virtual ADO from the kept state x [H][N][N] (the kernel's fixed-point rounds)."""
        U, N = self.count, self.N
        w = self.src_comm[:, None, None] * self._dv[self.src_mode] + self.src_anti[:, None, None] * self._sv[self.src_mode]
        contrib = w * x[self.src_ado]
        D = np.zeros((U, N, N), np.complex128)
        seg = np.repeat(np.arange(U), np.diff(self.src_offsets))
        np.add.at(D, seg, contrib)
        Q = self.evecs
        den = self.rate[:, None, None] - 1j * (self.evals[None, :, None] - self.evals[None, None, :])
        X = np.zeros_like(D)
        for _ in range(self.iterations):
            R = D + self._selfw[None] * X
            Rt = Q.conj().T[None] @ R @ Q[None]
            X = Q[None] @ (-Rt / den) @ Q.conj().T[None]
        return X

    def missing(self, x: np.ndarray) -> np.ndarray:
        """The terms the terminator adds, [H][N][N] (zero outside tier L)."""
        X = self.virtual(x)
        out = np.zeros_like(x)
        np.add.at(out, self.up_ado, -1j * self._dv[self.up_mode] * X[self.up_virtual])
        return out


def set_terminator(plan: Plan, term: "Terminator | None"):
    """Installs (or, with None, removes) the physics terminator on a plan (helix_hip_plan_set_terminator)."""
    if term is None:
        rc = hip().helix_hip_plan_set_terminator(plan._p, None)
    else:
        st = term.struct()
        rc = hip().helix_hip_plan_set_terminator(plan._p, C.byref(st))
    _check(rc, "helix_hip_plan_set_terminator")


def _sz_signs(spins: int):
    N = 1 << spins
    s = np.array([[1.0 if (i >> b) & 1 else -1.0 for i in range(N)] for b in range(spins)])  # [b][i]
    return s


def missing_terms(model: Model, state: np.ndarray, tier: int):
    """Tier-`tier` ADO indices of a deeper state and their terms c_n = -i sum_k [V_{b(k)}, rho_{n+e_k}]."""
    t = model.tables()
    rows = model.rows()
    spins = int(round(math.log2(model.N)))
    sg = _sz_signs(spins)
    idx = np.nonzero(rows.sum(1) == tier)[0]
    c = np.zeros((len(idx), model.N, model.N), np.complex128)
    for k in range(model.M):
        u = t["up"][idx, k]
        ok = u < model.H
        if not np.any(ok):
            continue
        b = t["mode_bath"][k]
        d = sg[b][:, None] - sg[b][None, :]
        c[ok] += -1j * d[None] * state[u[ok]]
    return idx, c


def element_bits(spins: int) -> np.ndarray:
    """[N*N][2*spins]: +-1 spin configuration of row i then column j for every element e = i*N + j."""
    N = 1 << spins
    e = np.arange(N * N)
    i, j = e // N, e % N
    bi = np.stack([np.where((i >> b) & 1, 1.0, -1.0) for b in range(spins)], 1)
    bj = np.stack([np.where((j >> b) & 1, 1.0, -1.0) for b in range(spins)], 1)
    return np.concatenate([bi, bj], 1).astype(np.float32)


def training_pairs(params_list, keep_tier: int = 2, deep_tier: int = 6, steps: int = 2000, every: int = 100,
                   dt: float = 0.01, min_weight: float = 1e-3, max_rate: float = 5.0):
    """Runs the plain FP64 solver at `deep_tier` for every parameter set (one batched launch per stage),
    records every `every`-th state, and returns (features, kappa*) rows, one per (set, tier-`keep_tier`
    ADO, element) whose block element was non-negligible over the window."""
    models = [p.model(deep_tier, "fp64", dt) for p in params_list]
    batch = capi.Batch(models, graph=True, autotune=False)
    batch.set_initial()
    num = [None] * len(models)
    den = [None] * len(models)
    for snap in range(steps // every):
        batch.step(every)
        for q, m in enumerate(models):
            state = batch.state(q)
            idx, c = missing_terms(m, state, keep_tier)
            rho = state[idx]
            if num[q] is None:
                num[q] = np.zeros_like(c)
                den[q] = np.zeros(c.shape)
            num[q] += np.conj(rho) * c
            den[q] += np.abs(rho) ** 2
    batch.close()
    X, Y = [], []
    for q, (p, m) in enumerate(zip(params_list, models)):
        rows = m.rows()
        idx = np.nonzero(rows.sum(1) == keep_tier)[0]
        extra = p.extra(rows[idx], keep_tier)
        bits = element_bits(p.spins)
        nn = num[q].reshape(len(idx), -1)
        dd = den[q].reshape(len(idx), -1)
        # elements that carry weight in the dynamics of their ADO (sum_t |rho_t|^2 >= min_weight x the largest)
        keep = dd > min_weight * dd.max(axis=1, keepdims=True)
        kappa = np.where(keep, nn / np.where(keep, dd, 1.0), 0.0)
        keep &= np.abs(kappa) < max_rate
        for a in range(len(idx)):
            sel = np.nonzero(keep[a])[0]
            if sel.size == 0:
                continue
            X.append(np.concatenate([bits[sel], np.tile(extra[a], (sel.size, 1))], 1))
            Y.append(np.stack([kappa[a, sel].real, kappa[a, sel].imag], 1))
    return np.concatenate(X).astype(np.float32), np.concatenate(Y).astype(np.float32)


class LearnedClosureSolver:
    """Truncated hierarchy (tiers 0..keep_tier, complex64 by default) plus the learned closure, with a
    calibration window against a deeper plain run that falls back to plain truncation when the closure's
    error exceeds `tolerance`."""

    def __init__(self, params: Params, mlp: MLP, keep_tier: int = 2, precision: str = "fp32", dt: float = 0.01):
        self.params, self.mlp, self.keep_tier, self.dt = params, mlp, keep_tier, dt
        self.model = params.model(keep_tier, precision, dt)
        self.plan = Plan(self.model, layout="full")  # the rates need not be conjugate-symmetric
        rows = self.model.rows()
        self.ados = np.ascontiguousarray(np.nonzero(rows.sum(1) == keep_tier)[0], np.int32)
        self.extra = np.ascontiguousarray(params.extra(rows[self.ados], keep_tier), np.float32)
        assert mlp.dims[0] == 2 * params.spins + self.extra.shape[1] and mlp.dims[3] == 2
        self.closure_on = False
        self.calibration_error = None

    def enable(self, on: bool = True):
        L = _bind()
        if on:
            rc = L.helix_hip_plan_set_closure(self.plan._p, self.mlp.handle,
                                              self.ados.ctypes.data_as(C.POINTER(C.c_int32)), len(self.ados),
                                              _f(self.extra), self.extra.shape[1])
        else:
            rc = L.helix_hip_plan_set_closure(self.plan._p, None, None, 0, None, 0)
        _check(rc, "helix_hip_plan_set_closure")
        self.closure_on = on

    def calibrate(self, window: int = 500, deep_tier: int = 6, tolerance: float | None = None) -> bool:
        """Runs `window` steps of the plain solver at `deep_tier` (the reference), of plain truncation and of
        the closure.  The closure is kept only when it is closer to the reference than plain truncation
        and, if `tolerance` is given, within max|d rho_S| <= tolerance; otherwise the solver falls back to
        plain truncation."""
        deep = Plan(self.params.model(deep_tier, "fp64", self.dt))
        deep.set_initial()
        deep.step(window)
        ref = deep.tier0()
        deep.close()
        self.enable(False)
        self.plan.set_initial()
        self.plan.step(window)
        self.plain_calibration_error = float(np.max(np.abs(self.plan.tier0() - ref)))
        self.enable(True)
        self.plan.set_initial()
        self.plan.step(window)
        self.calibration_error = float(np.max(np.abs(self.plan.tier0() - ref)))
        trusted = self.calibration_error < self.plain_calibration_error
        if tolerance is not None:
            trusted = trusted and self.calibration_error <= tolerance
        self.enable(trusted)
        self.plan.set_initial()
        return trusted

    def run(self, steps: int) -> np.ndarray:
        self.plan.set_initial()
        self.plan.step(steps)
        return self.plan.tier0()

    def close(self):
        self.plan.close()
