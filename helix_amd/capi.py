"""ctypes front of helix_amd's native libraries.

``libhelix_hip.so`` holds the gfx950 kernels behind the C-ABI of ``include/helix_hip.h``;
``libhelix.so`` holds the C++17 host set-up (``include/helix_setup.h``) and the public C++ API.
Nothing here computes physics: a missing library raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "libhelix_hip.so")
HOST_LIB = os.path.join(PKG, "libhelix.so")

FP64, FP32, SPLIT_FP32 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "split_fp32": SPLIT_FP32}
# helix_hip_layout: how each ADO block is stored on the device
LAYOUTS = {"auto": 0, "full": 1, "hermitian": 2}
LAYOUT_NAMES = {1: "full", 2: "hermitian"}

# every symbol include/helix_hip.h and include/helix_setup.h declare
HIP_SYMBOLS = [
    "helix_hip_plan_create", "helix_hip_set_initial", "helix_hip_set_state", "helix_hip_get_state", "helix_hip_rhs",
    "helix_hip_step", "helix_hip_get_tier0", "helix_hip_time_steps", "helix_hip_plan_info", "helix_hip_plan_destroy",
    "helix_hip_last_error", "helix_hip_batch_create", "helix_hip_batch_set_initial", "helix_hip_batch_step",
    "helix_hip_batch_get_tier0", "helix_hip_batch_get_state", "helix_hip_batch_time_steps", "helix_hip_batch_info",
    "helix_hip_batch_destroy", "helix_hip_plan_set_closure", "helix_hip_plan_set_closure_rates",
    "helix_hip_plan_layout", "helix_hip_batch_layout", "helix_hip_step_noalloc", "helix_hip_plan_set_terminator",
    "helix_hip_plan_set_terminator_gain", "helix_hip_get_ados", "helix_hip_batch_get_ados", "helix_hip_batch_launches",
    "helix_hip_measure_peaks",
]
SETUP_SYMBOLS = [
    "helix_model_legacy", "helix_model_spin_glass", "helix_model_spin_glass_random", "helix_model_sparse",
    "helix_model_use_dense_hamiltonian",
    "helix_model_problem",
    "helix_model_rows",
    "helix_model_energy_diagonal", "helix_model_energy_trace", "helix_model_free", "helix_setup_pade",
    "helix_setup_spin_glass_couplings", "helix_setup_hierarchy_size", "helix_setup_last_error",
]

IO_SYMBOLS = [
    "helix_io_write_energy_trace", "helix_io_hierarchy_hash", "helix_io_checkpoint_write", "helix_io_checkpoint_read",
    "helix_io_problem_hash", "helix_io_last_error",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SpinOperator(C.Structure):
    _fields_ = [("diag", _dp), ("nflips", C.c_int32), ("flip_masks", _ip), ("flip_coefs", _dp)]


class CheckpointInfo(C.Structure):
    """Mirror of helix_io_checkpoint."""
    _fields_ = [("n", C.c_int32), ("n_ado", C.c_int32), ("n_modes", C.c_int32), ("n_baths", C.c_int32),
                ("order", C.c_int32), ("precision", C.c_int32), ("steps", C.c_uint64), ("time", C.c_double),
                ("dt", C.c_double), ("hierarchy_hash", C.c_uint64), ("problem_hash", C.c_uint64)]


class Peaks(C.Structure):
    """helix_hip_peaks"""
    _fields_ = [("hbm_read_gbs", C.c_double), ("hbm_copy_gbs", C.c_double), ("mfma_f64_tflops", C.c_double),
                ("mfma_f32_tflops", C.c_double), ("bytes", C.c_size_t), ("compute_units", C.c_int)]


class Terminator(C.Structure):
    """Mirror of helix_hip_terminator."""
    _fields_ = [("count", C.c_int32), ("rate", _dp), ("src_offsets", _ip), ("src_ado", _ip), ("src_mode", _ip),
                ("src_comm", _dp), ("src_anti", _dp), ("n_up", C.c_int32), ("up_ado", _ip), ("up_mode", _ip),
                ("up_virtual", _ip), ("h_vectors", _dp), ("h_values", _dp), ("iterations", C.c_int32)]


class Problem(C.Structure):
    """Mirror of helix_hip_problem (read-only use from Python)."""
    _fields_ = [
        ("abi_version", C.c_int32), ("n", C.c_int32), ("n_ado", C.c_int32), ("n_modes", C.c_int32),
        ("n_baths", C.c_int32), ("up", _ip), ("down", _ip), ("mode_bath", _ip), ("tier", _ip),
        ("down_comm", _dp), ("down_anti", _dp), ("self_coef", _dp), ("xi", _dp), ("xi_counter", _dp),
        ("hamiltonian", SpinOperator), ("hamiltonian_dense", _dp), ("couplings", C.POINTER(SpinOperator)),
        ("precision", C.c_int32), ("order", C.c_int32), ("dt", C.c_double), ("layout", C.c_int32),
        ("hamiltonian_csr", C.c_void_p), ("couplings_csr", C.c_void_p),
    ]


_hip = None
_host = None


class HelixError(RuntimeError):
    pass


def _load(path):
    if not os.path.exists(path):
        raise HelixError(f"{os.path.basename(path)} is not built; run `python -m helix_amd.build`")
    return C.CDLL(path)


def hip():
    """libhelix_hip.so with argtypes bound."""
    global _hip
    if _hip is None:
        L = _load(HIP_LIB)
        vp = C.c_void_p
        L.helix_hip_plan_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.helix_hip_set_initial.argtypes = [vp, _dp]
        L.helix_hip_set_state.argtypes = [vp, _dp]
        L.helix_hip_get_state.argtypes = [vp, _dp]
        L.helix_hip_rhs.argtypes = [vp, _dp]
        L.helix_hip_step.argtypes = [vp, C.c_uint64, C.c_uint64, _dp, _dp]
        L.helix_hip_step_noalloc.argtypes = [vp, C.c_uint64, C.c_uint64, _dp, _dp]
        L.helix_hip_plan_set_terminator.argtypes = [vp, C.POINTER(Terminator)]
        L.helix_hip_get_tier0.argtypes = [vp, _dp]
        L.helix_hip_get_ados.argtypes = [vp, _ip, C.c_int, _dp]
        L.helix_hip_plan_set_terminator_gain.argtypes = [vp, _dp]
        L.helix_hip_time_steps.argtypes = [vp, C.c_uint64, _dp, _dp]
        L.helix_hip_plan_info.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.helix_hip_plan_destroy.argtypes = [vp]
        L.helix_hip_plan_destroy.restype = None
        L.helix_hip_last_error.restype = C.c_char_p
        L.helix_hip_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_uint, C.POINTER(vp)]
        L.helix_hip_batch_set_initial.argtypes = [vp, C.c_int, _dp]
        L.helix_hip_batch_step.argtypes = [vp, C.c_uint64]
        L.helix_hip_batch_get_tier0.argtypes = [vp, C.c_int, _dp]
        L.helix_hip_batch_get_state.argtypes = [vp, C.c_int, _dp]
        L.helix_hip_batch_get_ados.argtypes = [vp, C.c_int, _ip, C.c_int, _dp]
        L.helix_hip_batch_time_steps.argtypes = [vp, C.c_uint64, _dp]
        L.helix_hip_batch_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        L.helix_hip_batch_destroy.argtypes = [vp]
        L.helix_hip_batch_destroy.restype = None
        L.helix_hip_plan_layout.argtypes = [vp, _ip]
        L.helix_hip_batch_layout.argtypes = [vp, _ip]
        L.helix_hip_batch_launches.argtypes = [vp, _ip]
        L.helix_hip_measure_peaks.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(Peaks)]
        _hip = L
    return _hip


def host():
    """libhelix.so (set-up + public API) with argtypes bound."""
    global _host
    if _host is None:
        hip()  # resolve the kernel library first
        L = _load(HOST_LIB)
        vp = C.c_void_p
        L.helix_model_legacy.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.c_int, C.POINTER(vp)]
        L.helix_model_spin_glass.argtypes = [C.c_int, C.c_uint64, C.c_double, _dp, _dp, _dp, C.c_int, C.c_int,
                                             C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
        L.helix_model_use_dense_hamiltonian.argtypes = [vp, C.c_int]
        L.helix_model_sparse.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp,
                                         C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
        L.helix_model_spin_glass_random.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_int, C.POINTER(vp),
                                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.helix_model_problem.argtypes = [vp]
        L.helix_model_problem.restype = C.POINTER(Problem)
        L.helix_model_rows.argtypes = [vp, _ip]
        L.helix_model_energy_diagonal.argtypes = [vp, _dp]
        L.helix_model_energy_diagonal.restype = C.c_double
        L.helix_model_energy_trace.argtypes = [vp, _dp]
        L.helix_model_energy_trace.restype = C.c_double
        L.helix_model_free.argtypes = [vp]
        L.helix_model_free.restype = None
        L.helix_setup_pade.argtypes = [C.c_int, _dp, _dp]
        L.helix_setup_spin_glass_couplings.argtypes = [C.c_int, C.c_uint64, _dp]
        L.helix_setup_hierarchy_size.argtypes = [C.c_int, C.c_int]
        L.helix_setup_hierarchy_size.restype = C.c_long
        L.helix_setup_last_error.restype = C.c_char_p
        L.helix_io_write_energy_trace.argtypes = [C.c_char_p, _dp, _dp, C.c_int64, C.c_int]
        L.helix_io_hierarchy_hash.argtypes = [_ip, C.c_int64, C.c_int32]
        L.helix_io_hierarchy_hash.restype = C.c_uint64
        L.helix_io_checkpoint_write.argtypes = [C.c_char_p, C.POINTER(CheckpointInfo), _dp]
        L.helix_io_checkpoint_read.argtypes = [C.c_char_p, C.POINTER(CheckpointInfo), _dp, C.c_uint64]
        L.helix_io_problem_hash.argtypes = [C.c_void_p]
        L.helix_io_problem_hash.restype = C.c_uint64
        L.helix_io_last_error.restype = C.c_char_p
        _host = L
    return _host


def _pd(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _check(rc: int, what: str, lib=None):
    if rc != 0:
        msg = (lib or hip()).helix_hip_last_error() if lib is None else lib.helix_setup_last_error()
        raise HelixError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


class Model:
    """A tabulated HEOM problem built by the C++ host set-up."""

    def __init__(self, handle, N: int):
        self._h = handle
        self.N = N
        p = host().helix_model_problem(self._h).contents
        self.H = p.n_ado
        self.M = p.n_modes
        self.B = p.n_baths
        self.order = p.order
        self.dt = p.dt
        self.precision = p.precision

    @classmethod
    def legacy(cls, N=1024, beta=3.0, gamma=1.0, zeta=0.01, K=2, tiers=3, order=4, dt=0.1, precision="fp64"):
        h = C.c_void_p()
        rc = host().helix_model_legacy(N, beta, gamma, zeta, K, tiers, order, dt, PRECISIONS[precision], C.byref(h))
        _check(rc, "helix_model_legacy", host())
        return cls(h, N)

    @classmethod
    def spin_glass(cls, spins, seed, lam=0.1, gamma=1.0, T=1.0, K=1, L=4, h=1.0, order=4, dt=0.01,
                   precision="fp64", bath_params=None):
        if bath_params is None:
            bath_params = [(lam, gamma, T)] * spins
        lams = np.array([b[0] for b in bath_params], np.float64)
        gams = np.array([b[1] for b in bath_params], np.float64)
        temps = np.array([b[2] for b in bath_params], np.float64)
        hdl = C.c_void_p()
        rc = host().helix_model_spin_glass(spins, seed, h, _pd(lams), _pd(gams), _pd(temps), K, L, order, dt,
                                           PRECISIONS[precision], C.byref(hdl))
        _check(rc, "helix_model_spin_glass", host())
        return cls(hdl, 1 << spins)

    @classmethod
    def sparse(cls, hamiltonian, couplings, baths, K=1, max_tier=2, order=4, dt=0.01, precision="fp64"):
        """A general sparse system (System::from_sparse): Hermitian H and couplings given as dense arrays
        (their nonzeros are taken), one bath (beta, gamma, zeta) per coupling -- the reference's Bath
        fields, lambda = zeta gamma / 2.  Runs on the generic CSR kernels."""
        def csr(a):
            a = np.asarray(a, np.complex128)
            off, col, val = [0], [], []
            for i in range(a.shape[0]):
                nz = np.nonzero(a[i])[0]
                col.extend(nz.tolist())
                val.extend(a[i, nz].tolist())
                off.append(len(col))
            return (np.array(off, np.int32), np.array(col, np.int32),
                    np.ascontiguousarray(np.array(val, np.complex128)).view(np.float64))
        N = np.asarray(hamiltonian).shape[0]
        ho, hc, hv = csr(hamiltonian)
        parts = [csr(v) for v in couplings]
        v_nnz = np.array([len(c) for _, c, _ in parts], np.int32)
        v_off = np.ascontiguousarray(np.concatenate([o for o, _, _ in parts]), np.int32)
        v_col = np.ascontiguousarray(np.concatenate([c for _, c, _ in parts]), np.int32)
        v_val = np.ascontiguousarray(np.concatenate([v for _, _, v in parts]), np.float64)
        beta = np.array([b[0] for b in baths], np.float64)
        gamma = np.array([b[1] for b in baths], np.float64)
        zeta = np.array([b[2] for b in baths], np.float64)
        assert len(baths) == len(couplings)
        hdl = C.c_void_p()
        ip = lambda a: a.ctypes.data_as(_ip)  # noqa: E731
        _check(host().helix_model_sparse(N, len(hc), ip(ho), ip(hc), _pd(hv), len(couplings), ip(v_nnz), ip(v_off),
                                         ip(v_col), _pd(v_val), _pd(beta), _pd(gamma), _pd(zeta), K, max_tier, order,
                                         dt, PRECISIONS[precision], C.byref(hdl)), "helix_model_sparse", host())
        return cls(hdl, N)

    @classmethod
    def spin_glass_random(cls, seed, order=4, dt=0.01, precision="fp64"):
        """One instance of the batched workload (parameters drawn from SplitMix64(seed))."""
        hdl = C.c_void_p()
        sp, lt = C.c_int(), C.c_int()
        _check(host().helix_model_spin_glass_random(seed, order, dt, PRECISIONS[precision], C.byref(hdl), C.byref(sp),
                                                    C.byref(lt)), "helix_model_spin_glass_random", host())
        m = cls(hdl, 1 << sp.value)
        m.spins, m.max_tier = sp.value, lt.value
        return m

    def use_dense_hamiltonian(self, enable: bool = True):
        """Route -i[H, rho] through the dense MFMA commutator (H materialised as N x N)."""
        _check(host().helix_model_use_dense_hamiltonian(self._h, int(enable)), "use_dense_hamiltonian", host())
        return self

    def problem_ptr(self):
        return host().helix_model_problem(self._h)

    def hamiltonian_is_real(self) -> bool:
        """True when every Hamiltonian entry the problem carries is real (dense, CSR or spin-structured form)."""
        p = self.problem_ptr().contents
        N = self.N
        if p.hamiltonian_dense:
            h = np.ctypeslib.as_array(p.hamiltonian_dense, shape=(2 * N * N,))
            return bool(np.all(h[1::2] == 0.0))
        return True

    def problem_hash(self) -> int:
        """helix_io_problem_hash of this model's problem descriptor."""
        return int(host().helix_io_problem_hash(C.cast(self.problem_ptr(), C.c_void_p)))

    def rows(self) -> np.ndarray:
        r = np.zeros((self.H, self.M), np.int32)
        host().helix_model_rows(self._h, r.ctypes.data_as(_ip))
        return r

    def tables(self) -> dict:
        p = self.problem_ptr().contents
        H, M, B, N = self.H, self.M, self.B, self.N

        def arr(ptr, n, dt=np.float64):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).copy()
        out = dict(up=arr(p.up, H * M, np.int32).reshape(H, M), down=arr(p.down, H * M, np.int32).reshape(H, M),
                   mode_bath=arr(p.mode_bath, M, np.int32), tier=arr(p.tier, H, np.int32),
                   down_comm=arr(p.down_comm, 2 * H * M).view(np.complex128).reshape(H, M),
                   down_anti=arr(p.down_anti, 2 * H * M).view(np.complex128).reshape(H, M),
                   self_coef=arr(p.self_coef, 2 * H).view(np.complex128),
                   xi=arr(p.xi, 2 * B).view(np.complex128), xi_counter=arr(p.xi_counter, 2 * B).view(np.complex128),
                   h_diag=arr(p.hamiltonian.diag, N))
        nf = p.hamiltonian.nflips
        out["h_masks"] = arr(p.hamiltonian.flip_masks, nf, np.int32) if nf else np.zeros(0, np.int32)
        out["h_coefs"] = arr(p.hamiltonian.flip_coefs, nf) if nf else np.zeros(0)
        return out

    def energy_diagonal(self, rho0: np.ndarray) -> float:
        r = np.ascontiguousarray(rho0, np.complex128)
        return host().helix_model_energy_diagonal(self._h, _pd(r.view(np.float64)))

    def energy(self, rho0: np.ndarray) -> float:
        r = np.ascontiguousarray(rho0, np.complex128)
        return host().helix_model_energy_trace(self._h, _pd(r.view(np.float64)))

    def initial_rho(self) -> np.ndarray:
        r = np.zeros((self.N, self.N), np.complex128)
        r[self.N - 1, self.N - 1] = 1.0
        return r

    def close(self):
        if self._h:
            host().helix_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _with_layout(model, layout: str) -> Problem:
    """A copy of the model's problem descriptor with the storage layout set (the tables it points to stay
    owned by the model; a plan copies them at creation)."""
    desc = Problem()
    C.memmove(C.byref(desc), model.problem_ptr(), C.sizeof(Problem))
    desc.layout = LAYOUTS[layout]
    return desc


class Plan:
    """A device-resident HEOM propagator (helix_hip_plan)."""

    def __init__(self, model: Model, device: int = 0, layout: str = "auto"):
        self.model = model
        self.N, self.H = model.N, model.H
        self._p = C.c_void_p()
        desc = _with_layout(model, layout)
        _check(hip().helix_hip_plan_create(C.cast(C.pointer(desc), C.c_void_p), device, C.byref(self._p)),
               "helix_hip_plan_create")

    def layout(self) -> str:
        v = C.c_int32()
        _check(hip().helix_hip_plan_layout(self._p, C.byref(v)), "helix_hip_plan_layout")
        return LAYOUT_NAMES[v.value]

    def save_checkpoint(self, path: str, steps: int, time: float):
        m = self.model
        desc = m.problem_ptr().contents
        info = CheckpointInfo(n=m.N, n_ado=m.H, n_modes=m.M, n_baths=desc.n_baths, order=desc.order,
                              precision=desc.precision, steps=steps, time=time, dt=desc.dt,
                              hierarchy_hash=hierarchy_hash(m.rows()), problem_hash=m.problem_hash())
        checkpoint_write(path, info, self.get_state())

    def load_checkpoint(self, path: str):
        """Restores the state; returns (steps, time) stored with it."""
        m = self.model
        info, _ = checkpoint_read(path, with_state=False)
        if (info.n, info.n_ado, info.n_modes, info.hierarchy_hash) != (m.N, m.H, m.M, hierarchy_hash(m.rows())):
            raise HelixError("checkpoint belongs to another problem (Hilbert size or hierarchy differs)")
        desc = m.problem_ptr().contents
        if info.order != desc.order or info.dt != desc.dt:
            raise HelixError("checkpoint was written with another integration order or time step")
        if info.problem_hash != m.problem_hash():
            raise HelixError("checkpoint belongs to another problem (bath, Hamiltonian or coupling coefficients differ)")
        _, state = checkpoint_read(path)
        self.set_state(state)
        return info.steps, info.time

    def set_initial(self, rho0: np.ndarray | None = None):
        r = self.model.initial_rho() if rho0 is None else np.ascontiguousarray(rho0, np.complex128)
        _check(hip().helix_hip_set_initial(self._p, _pd(r.view(np.float64))), "helix_hip_set_initial")

    def set_state(self, state: np.ndarray):
        s = np.ascontiguousarray(state, np.complex128)
        assert s.shape == (self.H, self.N, self.N)
        _check(hip().helix_hip_set_state(self._p, _pd(s.view(np.float64))), "helix_hip_set_state")

    def get_state(self) -> np.ndarray:
        s = np.zeros((self.H, self.N, self.N), np.complex128)
        _check(hip().helix_hip_get_state(self._p, _pd(s.view(np.float64))), "helix_hip_get_state")
        return s

    def rhs(self) -> np.ndarray:
        s = np.zeros((self.H, self.N, self.N), np.complex128)
        _check(hip().helix_hip_rhs(self._p, _pd(s.view(np.float64))), "helix_hip_rhs")
        return s

    def step(self, nsteps: int, record_every: int = 0, record_rho: bool = False, record_diag: bool = False):
        nrec = nsteps // record_every if record_every else 0
        rho = np.zeros((nrec, self.N, self.N), np.complex128) if (nrec and record_rho) else None
        diag = np.zeros((nrec, self.N)) if (nrec and record_diag) else None
        _check(hip().helix_hip_step(self._p, nsteps, record_every,
                                    _pd(rho.view(np.float64)) if rho is not None else None,
                                    _pd(diag) if diag is not None else None), "helix_hip_step")
        return rho, diag

    def tier0(self) -> np.ndarray:
        r = np.zeros((self.N, self.N), np.complex128)
        _check(hip().helix_hip_get_tier0(self._p, _pd(r.view(np.float64))), "helix_hip_get_tier0")
        return r

    def ados(self, ados) -> np.ndarray:
        """Blocks of the listed ADOs (hierarchy order) as [len(ados)][N][N] complex128 (helix_hip_get_ados)."""
        idx = np.ascontiguousarray(ados, np.int32)
        out = np.zeros((idx.size, self.N, self.N), np.complex128)
        _check(hip().helix_hip_get_ados(self._p, idx.ctypes.data_as(_ip), idx.size, _pd(out.view(np.float64))),
               "helix_hip_get_ados")
        return out

    def time_steps(self, nsteps: int):
        st = np.zeros(self.model.order)
        tot = C.c_double()
        _check(hip().helix_hip_time_steps(self._p, nsteps, _pd(st), C.byref(tot)), "helix_hip_time_steps")
        return st, tot.value

    def info(self) -> dict:
        b, l, f, d = C.c_size_t(), C.c_int(), C.c_int(), C.c_int()
        _check(hip().helix_hip_plan_info(self._p, C.byref(b), C.byref(l), C.byref(f), C.byref(d)), "plan_info")
        return dict(device_bytes=b.value, launches_per_step=l.value, flip=bool(f.value), dense=bool(d.value))

    def close(self):
        if self._p:
            hip().helix_hip_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_GRAPH, BATCH_AUTOTUNE, BATCH_PER_SIZE, BATCH_PERSISTENT = 1, 2, 4, 8


class Batch:
    """Many independent problems advanced together: one launch per Taylor stage for all of them."""

    def __init__(self, models, device: int = 0, graph: bool = True, autotune: bool = True, layout: str = "auto",
                 per_size: bool = False, persistent: bool = False):
        self.models = list(models)
        descs = [_with_layout(m, layout) for m in self.models]
        arr = (C.c_void_p * len(descs))(*[C.cast(C.pointer(d), C.c_void_p) for d in descs])
        self._b = C.c_void_p()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_AUTOTUNE if autotune else 0) | (BATCH_PER_SIZE if per_size else 0)
        flags |= BATCH_PERSISTENT if persistent else 0
        _check(hip().helix_hip_batch_create(arr, len(self.models), device, flags, C.byref(self._b)),
               "helix_hip_batch_create")

    def layout(self) -> str:
        v = C.c_int32()
        _check(hip().helix_hip_batch_layout(self._b, C.byref(v)), "helix_hip_batch_layout")
        return LAYOUT_NAMES[v.value]

    def set_initial(self, index: int = -1, rho0: np.ndarray | None = None):
        r = None if rho0 is None else np.ascontiguousarray(rho0, np.complex128)
        _check(hip().helix_hip_batch_set_initial(self._b, index, _pd(r.view(np.float64)) if r is not None else None),
               "helix_hip_batch_set_initial")

    def step(self, nsteps: int):
        _check(hip().helix_hip_batch_step(self._b, nsteps), "helix_hip_batch_step")

    def tier0(self, index: int) -> np.ndarray:
        N = self.models[index].N
        r = np.zeros((N, N), np.complex128)
        _check(hip().helix_hip_batch_get_tier0(self._b, index, _pd(r.view(np.float64))), "batch_get_tier0")
        return r

    def state(self, index: int) -> np.ndarray:
        m = self.models[index]
        s = np.zeros((m.H, m.N, m.N), np.complex128)
        _check(hip().helix_hip_batch_get_state(self._b, index, _pd(s.view(np.float64))), "batch_get_state")
        return s

    def ados(self, index: int, ados) -> np.ndarray:
        """Blocks of the listed ADOs of problem `index` as [len(ados)][N][N] complex128 (helix_hip_batch_get_ados)."""
        N = self.models[index].N
        idx = np.ascontiguousarray(ados, np.int32)
        out = np.zeros((idx.size, N, N), np.complex128)
        _check(hip().helix_hip_batch_get_ados(self._b, index, idx.ctypes.data_as(_ip), idx.size,
                                              _pd(out.view(np.float64))), "helix_hip_batch_get_ados")
        return out

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(hip().helix_hip_batch_time_steps(self._b, nsteps, C.byref(ms)), "batch_time_steps")
        return ms.value

    def info(self) -> dict:
        a, b, t, g = C.c_int64(), C.c_size_t(), C.c_int(), C.c_int()
        _check(hip().helix_hip_batch_info(self._b, C.byref(a), C.byref(b), C.byref(t), C.byref(g)), "batch_info")
        n = C.c_int32()
        _check(hip().helix_hip_batch_launches(self._b, C.byref(n)), "batch_launches")
        return dict(total_ado=a.value, device_bytes=b.value, threads=t.value, graph=bool(g.value),
                    launches_per_step=n.value)

    def close(self):
        if self._b:
            hip().helix_hip_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def measure_peaks(device: int = 0, nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """Attainable peaks of the device (helix_hip_measure_peaks): HBM read and copy GB/s, FP64 and FP32 MFMA TFLOP/s."""
    out = Peaks()
    _check(hip().helix_hip_measure_peaks(device, nbytes, reps, C.byref(out)), "measure_peaks")
    return {"hbm_read_gbs": out.hbm_read_gbs, "hbm_copy_gbs": out.hbm_copy_gbs,
            "mfma_f64_tflops": out.mfma_f64_tflops, "mfma_f32_tflops": out.mfma_f32_tflops,
            "bytes": int(out.bytes), "compute_units": int(out.compute_units)}


def pade(K: int):
    p, r = np.zeros(K), np.zeros(K)
    _check(host().helix_setup_pade(K, _pd(p), _pd(r)), "helix_setup_pade", host())
    return p, r


def spin_glass_couplings(spins: int, seed: int) -> np.ndarray:
    J = np.zeros((spins, spins))
    _check(host().helix_setup_spin_glass_couplings(spins, seed, _pd(J)), "couplings", host())
    return J


def _io_check(rc: int, what: str):
    if rc != 0:
        msg = host().helix_io_last_error()
        raise HelixError(f"{what} failed: {msg.decode() if msg else ''}")


def write_energy_trace(path: str, times, energies, float32: bool = True):
    """The legacy CLI's outputEnergy.txt rows (include/helix_io.h)."""
    t = np.ascontiguousarray(times, np.float64)
    e = np.ascontiguousarray(energies, np.float64)
    assert t.shape == e.shape
    _io_check(host().helix_io_write_energy_trace(path.encode(), _pd(t), _pd(e), t.size, int(float32)),
              "helix_io_write_energy_trace")


def hierarchy_hash(rows: np.ndarray) -> int:
    r = np.ascontiguousarray(rows, np.int32)
    return int(host().helix_io_hierarchy_hash(r.ctypes.data_as(_ip), r.shape[0], r.shape[1]))


def checkpoint_write(path: str, info: CheckpointInfo, state: np.ndarray):
    s = np.ascontiguousarray(state, np.complex128)
    assert s.shape == (info.n_ado, info.n, info.n)
    _io_check(host().helix_io_checkpoint_write(path.encode(), C.byref(info), _pd(s.view(np.float64))),
              "helix_io_checkpoint_write")


def checkpoint_read(path: str, with_state: bool = True):
    """(info, state or None) of a checkpoint file; raises on any header or hash mismatch."""
    info = CheckpointInfo()
    _io_check(host().helix_io_checkpoint_read(path.encode(), C.byref(info), None, 0), "helix_io_checkpoint_read")
    if not with_state:
        return info, None
    s = np.zeros((info.n_ado, info.n, info.n), np.complex128)
    _io_check(host().helix_io_checkpoint_read(path.encode(), C.byref(info), _pd(s.view(np.float64)), s.size * 2),
              "helix_io_checkpoint_read")
    return info, s
