"""Python front of the CPU FP64 oracle (``libheom_oracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module.

Model builders here restate the reference's set-up independently of the
product's C++ host code:
  * legacy spin glass: src/initialize.cu:95-218, src/parameters.cu:11-39
  * seeded transverse-field spin glass with one sigma^z Drude-Lorentz bath per
    spin (the generalised model; bath parameters map as beta = 1/T,
    zeta = 2 lambda / gamma, omega0 = 1, cf. src/helix.cpp:375-384)
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libheom_oracle.so")
_lib = None

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("offsets", _ip), ("cols", _ip), ("vals", _dp)]


class _Problem(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("H", C.c_int), ("M", C.c_int), ("B", C.c_int),
        ("rows", _ip), ("edges", _ip), ("mode_bath", _ip), ("mode_kind", _ip),
        ("nu", _dp), ("kphi", _dp), ("kxi", _dp), ("kxic", _dp), ("kthc", _dp), ("ktha", _dp),
        ("h_diagonal", C.c_int), ("h_diag", _dp), ("h", _Csr), ("v", C.POINTER(_Csr)),
        ("threads", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.hx_psd.argtypes = [C.c_int, _dp, _dp]
        L.hx_hier_size.argtypes = [C.c_int, C.c_int]
        L.hx_hier_size.restype = C.c_long
        L.hx_hier_rows.argtypes = [C.c_int, C.c_int, _ip]
        L.hx_hier_rows.restype = C.c_long
        L.hx_hier_edges.argtypes = [_ip, C.c_int, C.c_int, _ip]
        L.hx_bath_coefficients.argtypes = [C.c_double] * 4 + [C.c_int, _dp]
        L.hx_legacy_operators.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                          _dp, _ip, _ip, _dp]
        L.hx_legacy_spin_count.argtypes = [C.c_int]
        L.hx_legacy_hamiltonian_diagonal.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
        L.hx_legacy_hamiltonian_diagonal.restype = C.c_double
        L.hx_legacy_coupling_matrix.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        L.hx_splitmix64.argtypes = [C.POINTER(C.c_ulonglong)]
        L.hx_splitmix64.restype = C.c_ulonglong
        L.hx_uniform.argtypes = [C.POINTER(C.c_ulonglong)]
        L.hx_uniform.restype = C.c_double
        L.hx_normal.argtypes = [C.POINTER(C.c_ulonglong)]
        L.hx_normal.restype = C.c_double
        L.hx_tfim_couplings.argtypes = [C.c_int, C.c_ulonglong, _dp]
        L.hx_tfim_hamiltonian.argtypes = [C.c_int, _dp, C.c_double, _ip, _ip, _dp]
        L.hx_rhs.argtypes = [C.POINTER(_Problem), _dp, _dp]
        L.hx_taylor_step.argtypes = [C.POINTER(_Problem), C.c_int, C.c_double, C.c_long, _dp]
        L.hx_rk4_step.argtypes = [C.POINTER(_Problem), C.c_double, C.c_long, _dp]
        L.hx_energy_diagonal.argtypes = [_dp, _dp, C.c_int, _dp]
        L.hx_energy_diagonal.restype = C.c_double
        L.hx_trace_h_rho.argtypes = [C.POINTER(_Csr), _dp, _dp]
        _lib = L
    return _lib


def _pi(a):
    return a.ctypes.data_as(_ip)


def _pd(a):
    return a.ctypes.data_as(_dp)


# ------------------------------------------------------------------ basics

def psd(K: int):
    poles = np.zeros(max(K, 1))
    res = np.zeros(max(K, 1))
    rc = lib().hx_psd(K, _pd(poles), _pd(res))
    if rc != 0:
        raise RuntimeError(f"psd failed rc={rc}")
    return poles[:K].copy(), res[:K].copy()


def hierarchy_size(modes: int, tiers: int) -> int:
    return int(lib().hx_hier_size(modes, tiers))


def hierarchy_rows(modes: int, tiers: int) -> np.ndarray:
    H = hierarchy_size(modes, tiers)
    rows = np.zeros((H, modes), dtype=np.int32)
    n = lib().hx_hier_rows(modes, tiers, _pi(rows))
    assert n == H
    return rows


def hierarchy_edges(rows: np.ndarray) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    H, M = rows.shape
    edges = np.zeros((H, 2 * M), dtype=np.int32)
    lib().hx_hier_edges(_pi(rows), H, M, _pi(edges))
    return edges


@dataclass
class BathCoefficients:
    nu: np.ndarray
    kxi: complex
    kxic: complex
    kthc: complex
    ktha: complex
    kphi: np.ndarray  # complex, K


def bath_coefficients(beta, gamma, zeta, omega0, K) -> BathCoefficients:
    out = np.zeros(K + 8 + 2 * K)
    rc = lib().hx_bath_coefficients(beta, gamma, zeta, omega0, K, _pd(out))
    if rc != 0:
        raise RuntimeError("bath coefficients failed")
    kphi = out[K + 8::2][:K] + 1j * out[K + 9::2][:K]
    return BathCoefficients(out[:K + 1].copy(), complex(out[K + 1], 0.0), complex(out[K + 2], out[K + 3]),
                            complex(out[K + 4], out[K + 5]), complex(out[K + 6], out[K + 7]), kphi)


@dataclass
class Csr:
    n: int
    offsets: np.ndarray
    cols: np.ndarray
    vals: np.ndarray  # complex128

    def dense(self) -> np.ndarray:
        d = np.zeros((self.n, self.n), dtype=np.complex128)
        for i in range(self.n):
            for z in range(self.offsets[i], self.offsets[i + 1]):
                d[i, self.cols[z]] += self.vals[z]
        return d

    @staticmethod
    def from_dense(d: np.ndarray) -> "Csr":
        n = d.shape[0]
        offs = [0]
        cols, vals = [], []
        for i in range(n):
            nz = np.nonzero(d[i])[0]
            cols.extend(nz.tolist())
            vals.extend(d[i, nz].tolist())
            offs.append(len(cols))
        return Csr(n, np.array(offs, np.int32), np.array(cols, np.int32), np.array(vals, np.complex128))

    def _struct(self):
        self._vals_iv = np.ascontiguousarray(self.vals.astype(np.complex128)).view(np.float64)
        return _Csr(self.n, len(self.cols), _pi(self.offsets), _pi(self.cols), _pd(self._vals_iv))


# ------------------------------------------------------------------ models

def legacy_operators(N=1024, omega0=1.0, dlong=1.0, dtrans=1.0, jn_aver=0.6, is_square=False):
    spins = lib().hx_legacy_spin_count(N)
    cap = N * (spins + 1)
    h = np.zeros(N)
    off = np.zeros(N + 1, np.int32)
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(2 * cap)
    nnz = lib().hx_legacy_operators(N, omega0, dlong, dtrans, jn_aver, int(is_square), _pd(h), _pi(off), _pi(cols),
                                    _pd(vals))
    v = Csr(N, off, cols[:nnz].copy(), vals[:2 * nnz].view(np.complex128).copy())
    return h, v


def tfim_couplings(spins: int, seed: int) -> np.ndarray:
    J = np.zeros((spins, spins))
    lib().hx_tfim_couplings(spins, C.c_ulonglong(seed), _pd(J))
    return J


def tfim_hamiltonian(spins: int, J: np.ndarray, h: float = 1.0) -> Csr:
    N = 1 << spins
    cap = N * (spins + 1)
    off = np.zeros(N + 1, np.int32)
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(2 * cap)
    J = np.ascontiguousarray(J, dtype=np.float64)
    nnz = lib().hx_tfim_hamiltonian(spins, _pd(J), h, _pi(off), _pi(cols), _pd(vals))
    return Csr(N, off, cols[:nnz].copy(), vals[:2 * nnz].view(np.complex128).copy())


def sigma_z(spins: int, spin: int) -> Csr:
    N = 1 << spins
    diag = np.where((np.arange(N) >> spin) & 1, 1.0, -1.0)
    return Csr(N, np.arange(N + 1, dtype=np.int32), np.arange(N, dtype=np.int32), diag.astype(np.complex128))


class Rng:
    """SplitMix64 stream (same generator as hx_splitmix64)."""

    def __init__(self, seed: int):
        self.state = C.c_ulonglong(seed)

    def uniform(self) -> float:
        return lib().hx_uniform(C.byref(self.state))

    def normal(self) -> float:
        return lib().hx_normal(C.byref(self.state))


@dataclass
class Bath:
    beta: float
    gamma: float
    zeta: float
    K: int
    V: Csr
    omega0: float = 1.0

    @staticmethod
    def drude_lorentz(lam: float, gamma: float, T: float, K: int, V: Csr) -> "Bath":
        return Bath(1.0 / T, gamma, 2.0 * lam / gamma, K, V)


@dataclass
class Problem:
    """One HEOM instance: operators, baths and hierarchy, in the oracle's layout."""
    N: int
    baths: list
    tiers: int
    h_diag: np.ndarray | None = None  # real diagonal (reference H_DIAGONAL path)
    h_csr: Csr | None = None
    threads: int = 0
    rows: np.ndarray = field(init=False)
    edges: np.ndarray = field(init=False)

    def __post_init__(self):
        self.B = len(self.baths)
        mb, mk, nu, kphi = [], [], [], []
        self.coefs = []
        for b, bath in enumerate(self.baths):
            c = bath_coefficients(bath.beta, bath.gamma, bath.zeta, bath.omega0, bath.K)
            self.coefs.append(c)
            for k in range(bath.K + 1):
                mb.append(b)
                mk.append(k)
                nu.append(c.nu[k])
                kphi.append(c.kphi[k - 1] if k > 0 else 0j)
        self.M = len(mb)
        self.mode_bath = np.array(mb, np.int32)
        self.mode_kind = np.array(mk, np.int32)
        self.nu = np.array(nu)
        self.kphi = np.array(kphi, np.complex128)
        self.kxi = np.array([c.kxi for c in self.coefs], np.complex128)
        self.kxic = np.array([c.kxic for c in self.coefs], np.complex128)
        self.kthc = np.array([c.kthc for c in self.coefs], np.complex128)
        self.ktha = np.array([c.ktha for c in self.coefs], np.complex128)
        self.rows = hierarchy_rows(self.M, self.tiers)
        self.edges = hierarchy_edges(self.rows)
        self.H = self.rows.shape[0]
        self._make_struct()

    def _make_struct(self):
        p = _Problem()
        p.N, p.H, p.M, p.B = self.N, self.H, self.M, self.B
        p.rows, p.edges = _pi(self.rows), _pi(self.edges)
        p.mode_bath, p.mode_kind = _pi(self.mode_bath), _pi(self.mode_kind)
        p.nu = _pd(self.nu)
        self._kphi_iv = self.kphi.view(np.float64)
        self._kxi_iv, self._kxic_iv = self.kxi.view(np.float64), self.kxic.view(np.float64)
        self._kthc_iv, self._ktha_iv = self.kthc.view(np.float64), self.ktha.view(np.float64)
        p.kphi, p.kxi, p.kxic = _pd(self._kphi_iv), _pd(self._kxi_iv), _pd(self._kxic_iv)
        p.kthc, p.ktha = _pd(self._kthc_iv), _pd(self._ktha_iv)
        if self.h_diag is not None:
            p.h_diagonal = 1
            self._hd = np.ascontiguousarray(np.asarray(self.h_diag, np.complex128)).view(np.float64)
            p.h_diag = _pd(self._hd)
        else:
            p.h_diagonal = 0
            p.h = self.h_csr._struct()
        self._vs = (_Csr * self.B)(*[bath.V._struct() for bath in self.baths])
        p.v = C.cast(self._vs, C.POINTER(_Csr))
        p.threads = self.threads
        self._p = p

    # ----- state helpers
    def zeros(self) -> np.ndarray:
        return np.zeros((self.H, self.N, self.N), np.complex128)

    def initial(self) -> np.ndarray:
        """rho_0 = |N-1><N-1| in tier 0 (src/initialize.cu:287-300)."""
        r = self.zeros()
        r[0, self.N - 1, self.N - 1] = 1.0
        return r

    def rhs(self, rho: np.ndarray) -> np.ndarray:
        rho = np.ascontiguousarray(rho, np.complex128)
        out = self.zeros()
        lib().hx_rhs(C.byref(self._p), _pd(rho.view(np.float64)), _pd(out.view(np.float64)))
        return out

    def taylor(self, rho: np.ndarray, dt: float, nsteps: int, order: int = 4) -> np.ndarray:
        rho = np.array(rho, np.complex128, order="C", copy=True)
        rc = lib().hx_taylor_step(C.byref(self._p), order, dt, nsteps, _pd(rho.view(np.float64)))
        if rc != 0:
            raise RuntimeError("taylor failed")
        return rho

    def rk4(self, rho: np.ndarray, dt: float, nsteps: int) -> np.ndarray:
        rho = np.array(rho, np.complex128, order="C", copy=True)
        lib().hx_rk4_step(C.byref(self._p), dt, nsteps, _pd(rho.view(np.float64)))
        return rho

    # ----- observables
    def h_dense(self) -> np.ndarray:
        if self.h_diag is not None:
            return np.diag(np.asarray(self.h_diag, np.complex128))
        return self.h_csr.dense()

    def energy_diagonal(self, rho: np.ndarray) -> float:
        """sum_i Re H_ii Re rho_ii (src/library/energy_observable.cpp:8-39)."""
        hd = np.ascontiguousarray(np.real(np.diag(self.h_dense())))
        r = np.ascontiguousarray(rho[0], np.complex128)
        return lib().hx_energy_diagonal(None, _pd(hd), self.N, _pd(r.view(np.float64)))

    def energy(self, rho: np.ndarray) -> float:
        """Re Tr(H rho_S)."""
        if self.h_csr is None:
            return self.energy_diagonal(rho)
        out = np.zeros(2)
        r = np.ascontiguousarray(rho[0], np.complex128)
        s = self.h_csr._struct()
        lib().hx_trace_h_rho(C.byref(s), _pd(r.view(np.float64)), _pd(out))
        return float(out[0])


# ------------------------------------------------------------- workloads

LEGACY = dict(N=1024, beta=3.0, gamma=1.0, zeta=0.01, K=2, tiers=3, dt=0.1, order=4)


def legacy_problem(N: int = 1024, beta=3.0, gamma=1.0, zeta=0.01, K=2, tiers=3, threads=0) -> Problem:
    """The reference's compiled legacy spin glass (src/parameters.cu:11-39)."""
    h, v = legacy_operators(N)
    return Problem(N, [Bath(beta, gamma, zeta, K, v)], tiers, h_diag=h, threads=threads)


def spin_glass_problem(spins: int, seed: int, lam=0.1, gamma=1.0, T=1.0, K=1, L=4, h=1.0,
                       bath_params=None, threads=0) -> Problem:
    """Seeded transverse-field spin glass, one sigma^z bath per spin, tiers 0..L."""
    J = tfim_couplings(spins, seed)
    H = tfim_hamiltonian(spins, J, h)
    baths = []
    for s in range(spins):
        lg, gg, tt = bath_params[s] if bath_params is not None else (lam, gamma, T)
        baths.append(Bath.drude_lorentz(lg, gg, tt, K, sigma_z(spins, s)))
    return Problem(1 << spins, baths, L + 1, h_csr=H, threads=threads)


def batched_instance_params(i: int):
    """Per-instance draw of the batched workload: seed 1000+i; spins, L, lambda, gamma, T in that order."""
    r = Rng(1000 + i)
    spins = 3 + min(int(r.uniform() * 4), 3)
    L = 2 + min(int(r.uniform() * 4), 3)
    lam = 0.05 + 0.45 * r.uniform()
    gamma = 0.5 + 1.5 * r.uniform()
    T = 0.5 + 1.5 * r.uniform()
    return dict(spins=spins, L=L, lam=lam, gamma=gamma, T=T, seed=1000 + i)


def normwise_rel(a: np.ndarray, b: np.ndarray) -> float:
    """max|a-b| / max|b| (the normwise relative error used for parity)."""
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def read_energy_trace(path: str):
    t, e = [], []
    with open(path) as f:
        for line in f:
            if line.strip():
                a, b = line.split()
                t.append(float(a))
                e.append(float(b))
    return np.array(t), np.array(e)
