"""Host-side mirror of the reference's operator interface over libsav_hip.so.

Names follow the reference: ``Config`` (src/Types.jl:27), ``FieldState``
(src/Types.jl:167), ``build_operators`` (src/SpectralUtils.jl:36),
``to_spectral!`` / ``to_physical!`` (src/SpectralUtils.jl:132-148),
``generate_initial_condition`` (src/Init.jl:74), ``run_simulation``
(src/simulation.jl:2).  Arrays use the Julia shapes and Fortran (column-major)
order, so ``state.phi.shape == (Nx, Ny, N)`` and the memory handed to the C ABI
is byte-identical to what Julia's ``ccall`` would pass.

Everything numeric runs in the HIP library; this module marshals arrays and prepares the
initial conditions of Init.jl on the host.
"""
from __future__ import annotations

import ctypes as ct
import math
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional

import numpy as np

from .config import Config, bdf_coeff, set_para_base, SplitMix64

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsav_hip.so"

SAV_OK = 0
SAV_W_KRYLOV_NOT_CONVERGED = 1
SAV_W_THETA2_NONNEG = 2
SAV_E_ARG = -1
SAV_E_HIP = -2
SAV_E_WC_NEGATIVE = -3
SAV_E_NONFINITE = -4
SAV_E_UNSUPPORTED = -5
SAV_PREC_FP64 = 0
SAV_PREC_FP32 = 1
SAV_PREC_MIXED = 2
SAV_PLANE_AUTO = 0
SAV_PLANE_REGISTER = 1
SAV_PLANE_LDS = 2
GRAPH_AUTO = 2  # Operators.set_graph: replay chosen by measurement
SAV_PRECOND_DIAG = 0
SAV_PRECOND_LEARNED = 1
SAV_PRECOND_STENCIL = 2
SAV_PRESENT = 0
SAV_OLD = 1
SAV_NET_NPARAM = 6249

_dp = ct.POINTER(ct.c_double)


class SavConfigC(ct.Structure):
    _fields_ = [("ndim", ct.c_int32), ("N", ct.c_int32), ("Nx", ct.c_int32), ("Ny", ct.c_int32),
                ("Nz", ct.c_int32), ("Lx", ct.c_double), ("Ly", ct.c_double), ("Lz", ct.c_double),
                ("epsilon", ct.c_double), ("M_phi", ct.c_double), ("M0_psi", ct.c_double),
                ("eta", ct.c_double), ("gamma_surf", ct.c_double), ("gamma_area", ct.c_int64),
                ("gamma_bend", ct.c_double), ("gamma_in", ct.c_int64), ("beta_in", ct.c_int64),
                ("gamma_out", ct.c_int64), ("beta_out", ct.c_int64), ("lamda", ct.c_int64),
                ("psi_in_v", _dp), ("psi_out_v", _dp),
                ("S1", ct.c_double), ("S2", ct.c_double), ("S3", ct.c_double), ("S4", ct.c_double),
                ("C1", ct.c_double), ("C2", ct.c_double), ("C3", ct.c_double),
                ("dt", ct.c_double), ("T", ct.c_double), ("tol", ct.c_double),
                ("precision", ct.c_int32), ("precond", ct.c_int32), ("itmax", ct.c_int32),
                ("atol", ct.c_double), ("device", ct.c_int32)]


class SavStepStatsC(ct.Structure):
    _fields_ = [("krylov_iters", ct.c_int32 * 3), ("krylov_solved", ct.c_int32 * 3),
                ("krylov_rnorm", ct.c_double * 3), ("energy", ct.c_double), ("area_ratio", ct.c_double),
                ("Q", ct.c_double), ("R1", ct.c_double), ("R2", ct.c_double), ("R3", ct.c_double),
                ("status", ct.c_int32)]


EXPORTS = ("sav_abi_version", "sav_create", "sav_destroy", "sav_last_error", "sav_set_state",
           "sav_get_state", "sav_init_sav", "sav_step", "sav_run", "sav_rfft", "sav_irfft", "sav_irfft_noalloc",
           "sav_fft_full", "sav_eval_functionals", "sav_psi_solve", "sav_modified_energy",
           "sav_last_elapsed_ms", "sav_kernel_timing", "sav_kernel_timing_get", "sav_set_precond",
           "sav_set_column_tile", "sav_autotune_info", "sav_column_tile_info", "sav_set_matvec_order", "sav_matvec_order_info", "sav_set_transform_tile", "sav_transform_tile_info", "sav_transform_tile_fwd_info", "sav_set_plane_kernel", "sav_plane_kernel_info", "sav_set_graph", "sav_graph_info", "sav_graph_choice", "sav_precond_init", "sav_precond_load", "sav_precond_save", "sav_precond_collect",
           "sav_precond_samples", "sav_precond_train", "sav_precond_train_target", "sav_net_forward", "sav_net_grad",
           "sav_ensemble_create", "sav_ensemble_destroy", "sav_ensemble_last_error", "sav_ensemble_size",
           "sav_ensemble_member", "sav_ensemble_run", "sav_ensemble_set_batched", "sav_ensemble_batched_info")

_lib_handle = None


class SavError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class WCNegativeError(SavError):
    """error("FATAL: W + C < 0") of src/Utils.jl:162,180,188."""


def load_library(path: Optional[os.PathLike] = None) -> ct.CDLL:
    """Loads libsav_hip.so; raises if it has not been built (no CPU fallback)."""
    global _lib_handle
    if _lib_handle is not None and path is None:
        return _lib_handle
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: build it with build.py (hipcc, gfx950)")
    lib = ct.CDLL(str(p))
    lib.sav_abi_version.restype = ct.c_int
    lib.sav_create.argtypes = [ct.POINTER(SavConfigC), ct.POINTER(ct.c_void_p)]
    lib.sav_destroy.argtypes = [ct.c_void_p]
    lib.sav_last_error.argtypes = [ct.c_void_p]
    lib.sav_last_error.restype = ct.c_char_p
    lib.sav_set_state.argtypes = [ct.c_void_p, ct.c_int] + [_dp] * 14 + [ct.c_double]
    lib.sav_get_state.argtypes = [ct.c_void_p, ct.c_int] + [_dp] * 14 + [_dp]
    lib.sav_init_sav.argtypes = [ct.c_void_p]
    lib.sav_step.argtypes = [ct.c_void_p, ct.c_double, ct.c_int, ct.POINTER(SavStepStatsC)]
    lib.sav_run.argtypes = [ct.c_void_p, _dp, ct.c_int, ct.c_int, _dp, _dp, ct.POINTER(ct.c_int32)]
    lib.sav_rfft.argtypes = [ct.c_void_p, _dp, _dp, ct.c_int]
    lib.sav_irfft.argtypes = [ct.c_void_p, _dp, _dp, ct.c_int]
    lib.sav_fft_full.argtypes = [ct.c_void_p, _dp, _dp, ct.c_int, ct.c_int]
    lib.sav_irfft_noalloc.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int]
    lib.sav_eval_functionals.argtypes = [ct.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.sav_psi_solve.argtypes = [ct.c_void_p, _dp, ct.c_double, ct.c_double, _dp, _dp,
                                  ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), _dp]
    lib.sav_modified_energy.argtypes = [ct.c_void_p, _dp, _dp]
    lib.sav_last_elapsed_ms.argtypes = [ct.c_void_p, _dp]
    lib.sav_kernel_timing.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_kernel_timing_get.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int64), _dp, _dp]
    _fp = ct.c_void_p
    lib.sav_set_precond.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_set_column_tile.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_autotune_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp]
    lib.sav_column_tile_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp, _dp]
    lib.sav_set_matvec_order.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_matvec_order_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp]
    lib.sav_set_transform_tile.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_transform_tile_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp, _dp]
    lib.sav_transform_tile_fwd_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp, _dp]
    lib.sav_set_plane_kernel.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_plane_kernel_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32), _dp, _dp]
    lib.sav_set_graph.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_graph_choice.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32)]
    lib.sav_graph_info.argtypes = [ct.c_void_p, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong),
                                   ct.POINTER(ct.c_longlong)]
    lib.sav_precond_init.argtypes = [ct.c_void_p, ct.c_uint64]
    lib.sav_precond_load.argtypes = [ct.c_void_p, _fp, ct.c_int]
    lib.sav_precond_save.argtypes = [ct.c_void_p, _fp, ct.c_int]
    lib.sav_precond_collect.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_uint64]
    lib.sav_precond_train_target.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_precond_samples.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]
    lib.sav_precond_train.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, _dp, _dp]
    lib.sav_net_forward.argtypes = [ct.c_void_p, _fp, ct.c_int, ct.c_int, ct.c_int, _fp]
    lib.sav_net_grad.argtypes = [ct.c_void_p, _fp, _fp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _fp, _dp]
    lib.sav_ensemble_create.argtypes = [ct.POINTER(SavConfigC), ct.c_int, ct.POINTER(ct.c_void_p)]
    lib.sav_ensemble_destroy.argtypes = [ct.c_void_p]
    lib.sav_ensemble_last_error.argtypes = [ct.c_void_p]
    lib.sav_ensemble_last_error.restype = ct.c_char_p
    lib.sav_ensemble_size.argtypes = [ct.c_void_p]
    lib.sav_ensemble_member.argtypes = [ct.c_void_p, ct.c_int]
    lib.sav_ensemble_member.restype = ct.c_void_p
    lib.sav_ensemble_run.argtypes = [ct.c_void_p, _dp, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _fp, _fp]
    lib.sav_ensemble_set_batched.argtypes = [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int32)]
    lib.sav_ensemble_batched_info.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
    if path is None:
        _lib_handle = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f(a, dtype):
    """Fortran-ordered contiguous array of the given dtype (no copy when possible)."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------------------
# FieldState (src/Types.jl:167-232) in Julia shapes, column-major
# ---------------------------------------------------------------------------
@dataclass
class FieldState:
    phi: np.ndarray
    phi_hat: np.ndarray
    psi: np.ndarray
    psi_hat: np.ndarray
    u: np.ndarray
    u_hat: np.ndarray
    p: np.ndarray
    p_hat: np.ndarray
    mu: np.ndarray
    mu_hat: np.ndarray
    nu: np.ndarray
    nu_hat: np.ndarray
    Q: float = 0.0
    R1: float = 0.0
    R2: float = 0.0
    R3: float = 0.0
    A0: np.ndarray = field(default_factory=lambda: np.zeros(1))
    dt: float = 0.0

    ARRAYS = ("phi", "phi_hat", "psi", "psi_hat", "u", "u_hat", "p", "p_hat", "mu", "mu_hat", "nu", "nu_hat")

    @staticmethod
    def zeros(conf: Config) -> "FieldState":
        """FieldState(Nx, Ny, N) of src/Types.jl:204-232 (3D: Types_3d.jl with 3 velocity components)."""
        sp = (conf.Nx, conf.Ny) + ((conf.Nz,) if conf.ndim == 3 else ())
        hp = (conf.Nxh,) + sp[1:]
        nd = conf.ndim
        r = lambda *s: np.zeros(s, dtype=np.float64, order="F")
        z = lambda *s: np.zeros(s, dtype=np.complex128, order="F")
        return FieldState(phi=r(*sp, conf.N), phi_hat=z(*hp, conf.N), psi=r(*sp, conf.N),
                          psi_hat=z(*hp, conf.N), u=r(*sp, nd), u_hat=z(*hp, nd), p=r(*sp),
                          p_hat=z(*hp), mu=r(*sp, conf.N), mu_hat=z(*hp, conf.N), nu=r(*sp, conf.N),
                          nu_hat=z(*hp, conf.N), A0=np.zeros(conf.N))

    def copy(self) -> "FieldState":
        return FieldState(**{k: (v.copy(order="K") if isinstance(v, np.ndarray) else v)
                             for k, v in self.__dict__.items()})


def update_state(dest: FieldState, src: FieldState) -> FieldState:
    """update_state! (src/Types.jl:243-265): copies the 12 arrays and Q, R1..R3; not A0, not dt."""
    for k in FieldState.ARRAYS:
        getattr(dest, k)[...] = getattr(src, k)
    dest.Q, dest.R1, dest.R2, dest.R3 = src.Q, src.R1, src.R2, src.R3
    return dest


def config_c(conf: Config, precision: int, itmax: int = 20, atol: float = 1e-14, device: int = 0,
             precond: int = SAV_PRECOND_DIAG):
    """sav_config_t of a Config (src/Types.jl:27-74); returns it with the arrays it points to."""
    pin = np.ascontiguousarray(conf.psi_in_v[: conf.N], dtype=np.float64)
    pout = np.ascontiguousarray(conf.psi_out_v[: conf.N], dtype=np.float64)
    c = SavConfigC()
    c.ndim, c.N, c.Nx, c.Ny, c.Nz = conf.ndim, conf.N, conf.Nx, conf.Ny, conf.Nz
    c.Lx, c.Ly, c.Lz = conf.Lx, conf.Ly, conf.Lz
    c.epsilon, c.M_phi, c.M0_psi, c.eta, c.gamma_surf = (conf.epsilon, conf.M_phi, conf.M0_psi,
                                                        conf.eta, conf.gamma_surf)
    c.gamma_area, c.gamma_bend = conf.gamma_area, conf.gamma_bend
    c.gamma_in, c.beta_in, c.gamma_out, c.beta_out, c.lamda = (conf.gamma_in, conf.beta_in,
                                                               conf.gamma_out, conf.beta_out, conf.lamda)
    c.psi_in_v, c.psi_out_v = _ptr(pin), _ptr(pout)
    c.S1, c.S2, c.S3, c.S4 = conf.S1, conf.S2, conf.S3, conf.S4
    c.C1, c.C2, c.C3 = conf.C1, conf.C2, conf.C3
    c.dt, c.T, c.tol = conf.dt, conf.T, conf.tol
    c.precision, c.precond, c.itmax, c.atol, c.device = precision, precond, itmax, atol, device
    return c, (pin, pout)


# ---------------------------------------------------------------------------
# Operators: the device handle (build_operators, src/SpectralUtils.jl:36-118)
# ---------------------------------------------------------------------------
class Operators:
    """Owns one libsav_hip handle: plans, twiddles, multipliers, device state.

    ``precision`` is SAV_PREC_FP64 (the reference's arithmetic), SAV_PREC_FP32
    (FP32 fields/FFTs with FP64 reductions, 3x3 solve and Krylov scalars) or
    SAV_PREC_MIXED (FP64 throughout except the inner iterations of the 2D psi solve,
    an FP32 BiCGSTAB inside an FP64 defect correction with the reference's stopping test)."""

    def __init__(self, conf: Config, precision: int = SAV_PREC_FP64, itmax: int = 20, atol: float = 1e-14,
                 device: int = 0, precond: int = SAV_PRECOND_DIAG, _borrow=None):
        self.lib = load_library()
        self.conf = conf
        self.precision = precision
        self.device = device
        self._owned = _borrow is None
        if _borrow is not None:
            self.h = _borrow
            return
        c, self._keep = config_c(conf, precision, itmax, atol, device, precond)
        h = ct.c_void_p()
        rc = self.lib.sav_create(ct.byref(c), ct.byref(h))
        if rc != 0:
            raise SavError(rc, self.lib.sav_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self.lib.sav_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            msg = self.lib.sav_last_error(self.h).decode()
            if rc == SAV_E_WC_NEGATIVE:
                raise WCNegativeError(rc, f"{what}: {msg}")
            raise SavError(rc, f"{what}: {msg}")
        return rc

    # -- state ----------------------------------------------------------------
    def set_state(self, which: int, st: FieldState):
        arr = []
        keep = []
        for k in FieldState.ARRAYS:
            v = getattr(st, k)
            a = _f(v, np.complex128 if k.endswith("_hat") else np.float64)
            keep.append(a)
            arr.append(_ptr(a))
        scal = np.array([st.Q, st.R1, st.R2, st.R3], dtype=np.float64)
        A0 = np.ascontiguousarray(st.A0, dtype=np.float64)
        if A0.size < self.conf.N:
            A0 = np.zeros(self.conf.N)
        self._check(self.lib.sav_set_state(self.h, which, *arr, _ptr(scal), _ptr(A0), float(st.dt)), "set_state")

    def get_state(self, which: int, st: Optional[FieldState] = None) -> FieldState:
        st = st or FieldState.zeros(self.conf)
        ptrs = []
        for k in FieldState.ARRAYS:
            a = getattr(st, k)
            assert a.flags.f_contiguous
            ptrs.append(_ptr(a))
        scal = np.zeros(4)
        A0 = np.zeros(self.conf.N)
        dt = ct.c_double(0.0)
        self._check(self.lib.sav_get_state(self.h, which, *ptrs, _ptr(scal), _ptr(A0), ct.byref(dt)), "get_state")
        st.Q, st.R1, st.R2, st.R3 = (float(x) for x in scal)
        if which == SAV_PRESENT:
            st.A0 = A0
        st.dt = dt.value
        return st

    def init_sav(self):
        """simulation.jl:24-53 on the present state (old = present afterwards)."""
        return self._check(self.lib.sav_init_sav(self.h), "init_sav")

    def step(self, dt: float, n: int):
        """One loop body (simulation.jl:73-110) + monitor; returns the step stats."""
        st = SavStepStatsC()
        rc = self._check(self.lib.sav_step(self.h, float(dt), int(n), ct.byref(st)), "step")
        return rc, st

    def run(self, t_nodes: np.ndarray, n0: int = 1):
        t = np.ascontiguousarray(t_nodes, dtype=np.float64)
        m = max(len(t) - 1, 0)
        E = np.zeros(m)
        A = np.zeros(m)
        kit = ct.c_int32(0)
        rc = self._check(self.lib.sav_run(self.h, _ptr(t), len(t), n0, _ptr(E), _ptr(A), ct.byref(kit)), "run")
        return rc, E, A, kit.value

    def elapsed_ms(self) -> float:
        v = ct.c_double(0.0)
        self.lib.sav_last_elapsed_ms(self.h, ct.byref(v))
        return v.value

    KERNELS = {1: "k_mv_cols_inv", 2: "k_mv_rows", 3: "k_mv_cols_fwd"}

    def kernel_timing(self, kernel_id: int):
        """Starts HIP-event timing of one psi-matvec kernel (0 stops)."""
        self._check(self.lib.sav_kernel_timing(self.h, int(kernel_id)), "kernel_timing")

    def kernel_timing_get(self):
        n, ms, b = ct.c_int64(0), ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_kernel_timing_get(self.h, ct.byref(n), ct.byref(ms), ct.byref(b)),
                    "kernel_timing_get")
        return n.value, ms.value, b.value

    # -- transforms (to_spectral!, to_physical!) --------------------------------
    def _count(self, a, half=False):
        sp = self.conf.Nx * self.conf.Ny * (self.conf.Nz if self.conf.ndim == 3 else 1)
        if half:
            sp = sp // self.conf.Nx * self.conf.Nxh
        return a.size // sp

    def rfft(self, u: np.ndarray) -> np.ndarray:
        u = _f(u, np.float64)
        cnt = self._count(u)
        shp = (self.conf.Nxh,) + u.shape[1:]
        out = np.zeros(shp, dtype=np.complex128, order="F")
        self._check(self.lib.sav_rfft(self.h, _ptr(u), _ptr(out), cnt), "rfft")
        return out

    def rfft_accurate(self, u: np.ndarray) -> np.ndarray:
        """Spectra of initial data in FP64: this handle's transform for an FP64 or mixed handle;
        for an FP32 handle the transform of a temporary FP64 handle of the same configuration
        (the FP32 FFT roundoff floor at high k would be amplified by the S1 k^4 stabiliser).
        Used to prepare inputs only, never inside a time step."""
        if self.precision in (SAV_PREC_FP64, SAV_PREC_MIXED):
            return self.rfft(u)
        h = Operators(self.conf, precision=SAV_PREC_FP64, device=self.device)
        try:
            return h.rfft(u)
        finally:
            h.close()

    def irfft(self, uh: np.ndarray) -> np.ndarray:
        uh = _f(uh, np.complex128)
        cnt = self._count(uh, half=True)
        shp = (self.conf.Nx,) + uh.shape[1:]
        out = np.zeros(shp, dtype=np.float64, order="F")
        self._check(self.lib.sav_irfft(self.h, _ptr(uh), _ptr(out), cnt), "irfft")
        return out

    def fft_full(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        x = _f(x, np.complex128)
        cnt = self._count(x)
        out = np.zeros_like(x, order="F")
        self._check(self.lib.sav_fft_full(self.h, _ptr(x), _ptr(out), cnt,
                                          int(inverse)), "fft_full")
        return out

    def eval_functionals(self, phi, psi):
        phi = _f(phi, np.float64)
        psi = _f(psi, np.float64)
        H1, H2, H3, G = (np.zeros_like(phi, order="F") for _ in range(4))
        W = np.zeros(3)
        area = np.zeros(self.conf.N)
        self._check(self.lib.sav_eval_functionals(self.h, _ptr(phi), _ptr(psi), _ptr(H1), _ptr(H2), _ptr(H3),
                                                  _ptr(G), _ptr(W), _ptr(area)), "eval_functionals")
        return H1, H2, H3, G, W, area

    def psi_solve(self, phi_star, a, dt, rhs_hat):
        phi_star = _f(phi_star, np.float64)
        rhs = _f(rhs_hat, np.complex128)
        out = np.zeros_like(rhs, order="F")
        it, sv, rn = ct.c_int32(0), ct.c_int32(0), ct.c_double(0)
        rc = self._check(self.lib.sav_psi_solve(self.h, _ptr(phi_star), float(a), float(dt),
                                                _ptr(rhs), _ptr(out),
                                                ct.byref(it), ct.byref(sv), ct.byref(rn)), "psi_solve")
        return out, dict(iters=it.value, solved=bool(sv.value), rnorm=rn.value, rc=rc)

    def modified_energy(self):
        E, ar = ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_modified_energy(self.h, ct.byref(E), ct.byref(ar)), "modified_energy")
        return E.value, ar.value

    # -- psi matvec autotuner (no reference counterpart) -------------------------
    def set_column_tile(self, columns: int):
        """4, 8 or 16 columns per workgroup in the psi matvec's column kernels; 0 retunes."""
        self._check(self.lib.sav_set_column_tile(self.h, int(columns)), "set_column_tile")

    def autotune_info(self):
        cw, a, b, c = ct.c_int32(0), ct.c_double(0), ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_column_tile_info(self.h, ct.byref(cw), ct.byref(c), ct.byref(a), ct.byref(b)),
                    "autotune_info")
        return dict(columns=cw.value, ms_4=c.value, ms_8=a.value, ms_16=b.value)

    def set_matvec_order(self, order: int):
        """Order of the psi matvec's kernels over a batch of systems: 1 all systems per launch,
        2 system by system over one shared intermediate pair, 0 retunes at the next solve."""
        self._check(self.lib.sav_set_matvec_order(self.h, int(order)), "set_matvec_order")

    def matvec_order_info(self):
        o, a, b = ct.c_int32(0), ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_matvec_order_info(self.h, ct.byref(o), ct.byref(a), ct.byref(b)),
                    "matvec_order_info")
        return dict(order=o.value, ms_batched=a.value, ms_sequenced=b.value)

    # -- graph replay of steps (no reference counterpart) -------------------------
    def set_transform_tile(self, columns: int):
        """Columns per workgroup of the column transforms outside the psi matvec (4, 8, 16; 0 retunes)."""
        return self._check(self.lib.sav_set_transform_tile(self.h, int(columns)), "set_transform_tile")

    def transform_tile_info(self):
        cw = ct.c_int32(0)
        m4, m8, m16 = ct.c_double(0), ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_transform_tile_info(self.h, ct.byref(cw), ct.byref(m4), ct.byref(m8), ct.byref(m16)),
                    "transform_tile_info")
        return dict(columns=cw.value, ms_4=m4.value, ms_8=m8.value, ms_16=m16.value)

    def transform_tile_fwd_info(self):
        cw = ct.c_int32(0)
        m4, m8, m16 = ct.c_double(0), ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_transform_tile_fwd_info(self.h, ct.byref(cw), ct.byref(m4), ct.byref(m8), ct.byref(m16)),
                    "transform_tile_fwd_info")
        return dict(columns=cw.value, ms_4=m4.value, ms_8=m8.value, ms_16=m16.value)

    def set_plane_kernel(self, kind: int):
        """(x, y) plane kernels of 128 x 128 planes in 3D: SAV_PLANE_AUTO (time both at the
        next step), SAV_PLANE_REGISTER or SAV_PLANE_LDS."""
        return self._check(self.lib.sav_set_plane_kernel(self.h, int(kind)), "set_plane_kernel")

    def plane_kernel_info(self):
        k = ct.c_int32(0)
        a, b = ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_plane_kernel_info(self.h, ct.byref(k), ct.byref(a), ct.byref(b)),
                    "plane_kernel_info")
        return dict(kind=k.value, ms_register=a.value, ms_lds=b.value)

    def set_graph(self, on):
        """Replay each step as one captured HIP graph (bit-identical results): False / 0 off,
        True / 1 on, GRAPH_AUTO (2) chosen by timing eager against replayed steps."""
        self._check(self.lib.sav_set_graph(self.h, int(on)), "set_graph")

    def graph_info(self):
        a, b, c = ct.c_longlong(0), ct.c_longlong(0), ct.c_longlong(0)
        self._check(self.lib.sav_graph_info(self.h, ct.byref(a), ct.byref(b), ct.byref(c)), "graph_info")
        ch = ct.c_int32(0)
        self._check(self.lib.sav_graph_choice(self.h, ct.byref(ch)), "graph_choice")
        return dict(replays=a.value, repeats=b.value, graphs=c.value, choice=ch.value)

    # -- learned psi preconditioner (no reference counterpart) -----------------
    def set_precond(self, precond: int):
        """SAV_PRECOND_DIAG (Solvers.jl:57) or SAV_PRECOND_LEARNED for the following psi solves."""
        self._check(self.lib.sav_set_precond(self.h, int(precond)), "set_precond")

    def precond_init(self, seed: int):
        self._check(self.lib.sav_precond_init(self.h, int(seed) & (2 ** 64 - 1)), "precond_init")

    def precond_load(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != SAV_NET_NPARAM:
            raise ValueError(f"expected {SAV_NET_NPARAM} parameters, got {p.size}")
        self._check(self.lib.sav_precond_load(self.h, p.ctypes.data, p.size), "precond_load")

    def precond_save(self) -> np.ndarray:
        p = np.zeros(SAV_NET_NPARAM, dtype=np.float32)
        self._check(self.lib.sav_precond_save(self.h, p.ctypes.data, p.size), "precond_save")
        return p

    def precond_collect(self, on: bool, patches_per_snapshot: int = 16, seed: int = 1):
        """Log training windows from the following (diagonal) psi solves."""
        self._check(self.lib.sav_precond_collect(self.h, int(bool(on)), int(patches_per_snapshot),
                                                 int(seed) & (2 ** 64 - 1)), "precond_collect")

    def precond_samples(self) -> int:
        n = ct.c_int64(0)
        self._check(self.lib.sav_precond_samples(self.h, ct.byref(n), None), "precond_samples")
        return n.value

    def precond_fallbacks(self) -> int:
        """Learned/stencil solves repeated with the diagonal preconditioner (no convergence)."""
        n = ct.c_int64(0)
        self._check(self.lib.sav_precond_samples(self.h, None, ct.byref(n)), "precond_samples")
        return n.value

    def precond_train_target(self, target: int):
        """Targets of the following training: 0 the solver's error corrections, 1 the analytic
        first-order stencil's output on the same windows (distillation)."""
        self._check(self.lib.sav_precond_train_target(self.h, int(target)), "precond_train_target")

    def precond_train(self, iters: int = 1000, batch: int = 32, lr: float = 2e-3, seed: int = 7):
        l0, l1 = ct.c_double(0), ct.c_double(0)
        self._check(self.lib.sav_precond_train(self.h, int(iters), int(batch), float(lr), int(seed) & (2 ** 64 - 1),
                                               ct.byref(l0), ct.byref(l1)), "precond_train")
        return l0.value, l1.value

    def net_forward(self, x: np.ndarray) -> np.ndarray:
        """x[nimg, H, W, 2] (float, rounded to BF16 on the way in) -> stencil weights y[nimg, H, W, 9]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        nimg, H, W, two = x.shape
        assert two == 2
        y = np.zeros((nimg, H, W, 9), dtype=np.float32)
        self._check(self.lib.sav_net_forward(self.h, x.ctypes.data, nimg, H, W, y.ctypes.data), "net_forward")
        return y

    def net_grad(self, x: np.ndarray, target: np.ndarray, border: int):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        nimg, H, W, _ = x.shape
        g = np.zeros(SAV_NET_NPARAM, dtype=np.float32)
        loss = ct.c_double(0)
        self._check(self.lib.sav_net_grad(self.h, x.ctypes.data, t.ctypes.data, nimg, H, W, int(border),
                                          g.ctypes.data, ct.byref(loss)), "net_grad")
        return loss.value, g


class Ensemble:
    """Independent instances driven concurrently (sav_ensemble_*).  ``members[i]`` are
    borrowed Operators for state set-up and read-back."""

    def __init__(self, confs: List[Config], precision: int = SAV_PREC_FP64, itmax: int = 20, atol: float = 1e-14,
                 device: int = 0):
        self.lib = load_library()
        self.confs = list(confs)
        arr = (SavConfigC * len(self.confs))()
        self._keep = []
        for i, conf in enumerate(self.confs):
            c, keep = config_c(conf, precision, itmax, atol, device)
            arr[i] = c
            self._keep.append(keep)
        e = ct.c_void_p()
        rc = self.lib.sav_ensemble_create(arr, len(self.confs), ct.byref(e))
        if rc != 0:
            raise SavError(rc, self.lib.sav_ensemble_last_error(None).decode())
        self.e = e
        self.members = [Operators(conf, precision, _borrow=ct.c_void_p(self.lib.sav_ensemble_member(e, i)))
                        for i, conf in enumerate(self.confs)]

    def set_graph(self, on: bool):
        """Graph replay of steps on every member (Operators.set_graph)."""
        for m in self.members:
            m.set_graph(on)

    def set_batched(self, on: bool) -> int:
        """Size classes: members with the same grid, N and solve settings run their psi solves
        as one batched solve per step (sav_ensemble_set_batched); returns the group count."""
        ng = ct.c_int32(0)
        rc = self.lib.sav_ensemble_set_batched(self.e, int(bool(on)), ct.byref(ng))
        if rc != 0:
            raise SavError(rc, "ensemble_set_batched: " + self.lib.sav_ensemble_last_error(self.e).decode())
        return ng.value

    def batched_info(self):
        """(group of each member (-1: none), batched solves run by each member's group)."""
        n = len(self.confs)
        g = np.zeros(n, dtype=np.int32)
        ns = np.zeros(n, dtype=np.int64)
        self.lib.sav_ensemble_batched_info(self.e, g.ctypes.data, ns.ctypes.data)
        return g, ns

    def run(self, t_nodes: np.ndarray, n0: int = 1, threads: int = 8):
        t = np.ascontiguousarray(t_nodes, dtype=np.float64)
        n, m = len(self.confs), max(len(t) - 1, 0)
        E = np.zeros((n, m))
        A = np.zeros((n, m))
        kit = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        rc = self.lib.sav_ensemble_run(self.e, _ptr(t), len(t), int(n0), int(threads), _ptr(E), _ptr(A),
                                       kit.ctypes.data, st.ctypes.data)
        if rc < 0:
            raise SavError(rc, "ensemble_run: " + self.lib.sav_ensemble_last_error(self.e).decode())
        return rc, E, A, kit, st

    def close(self):
        if getattr(self, "e", None):
            for m in self.members:
                m.h = None
            self.lib.sav_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_operators(conf: Config, **kw) -> Operators:
    """build_operators(cfg) of src/SpectralUtils.jl:36 -> a device handle."""
    return Operators(conf, **kw)


def to_spectral(u: np.ndarray, ops: Operators) -> np.ndarray:
    """to_spectral! (src/SpectralUtils.jl:132-135): unnormalised rfft over (x, y[, z])."""
    return ops.rfft(u)


def to_physical(u_hat: np.ndarray, ops: Operators) -> np.ndarray:
    """to_physical! (src/SpectralUtils.jl:144-148): irfft scaled by 1/(Nx Ny[Nz]); input untouched."""
    return ops.irfft(u_hat)


# ---------------------------------------------------------------------------
# Initial conditions (src/Init.jl:74-217), with the unseeded rand() of
# state_type 4 replaced by a seeded SplitMix64 stream.
# ---------------------------------------------------------------------------
def initial_fields_3d(conf: Config, state_type: int):
    """(phi, psi) in Julia shape (Nx, Ny, Nz, N) of Init_3d.jl:67-153.  N = 1, state_type 1: the ellipsoid
    tanh((0.2 Lx - sqrt((x-cx)^2/2 + (y-cy)^2/2 + (z-cz)^2)) / (sqrt 2 eps)).
    N = 2, state_type 7: two star shapes centred at x = 0.73 and 1.27.  The reference's
    branch indexes 2D grids (X, Y) and a 3D slice phi[:, :, n] of a 4D array, which Julia
    rejects; here the same formula R0 + amp cos(k theta) - r is evaluated with the 3D
    distance r and the azimuth theta = atan(x - cx, y - cy)."""
    Nx, Ny, Nz, N = conf.Nx, conf.Ny, conf.Nz, conf.N
    Lx, Ly, Lz, eps = conf.Lx, conf.Ly, conf.Lz, conf.epsilon
    x = np.linspace(0, Lx - Lx / Nx, Nx)[:, None, None]
    y = np.linspace(0, Ly - Ly / Ny, Ny)[None, :, None]
    z = np.linspace(0, Lz - Lz / Nz, Nz)[None, None, :]
    s2e = math.sqrt(2) * eps
    phi = np.zeros((Nx, Ny, Nz, N), order="F")
    # Init_3d.jl:96-123: the branches that the reference leaves without a shape (N = 1 with
    # state_type <= 6 other than 1, N = 2 with state_type >= 8, N >= 3) keep phi = 0
    if N == 1:
        if state_type > 6:
            raise ValueError(f"state_type {state_type} invalid for N = 1 (Init_3d.jl:98)")
        if state_type == 1:
            R = 0.2 * Lx
            dist = np.sqrt((x - Lx / 2) ** 2 / 2 + (y - Ly / 2) ** 2 / 2 + (z - Lz / 2) ** 2)
            phi[..., 0] = np.tanh((R - dist) / s2e)
    elif N == 2:
        if state_type <= 6:
            raise ValueError(f"state_type {state_type} invalid for N = 2 (Init_3d.jl:108)")
        if state_type == 7:
            if conf.goal == "s":
                raise ValueError("state_type 7 needs goal g (Init_3d.jl:114)")
            for n, cxn in enumerate((0.73, 1.27)):
                r = np.sqrt((x - cxn) ** 2 + (y - Ly / 2) ** 2 + (z - Lz / 2) ** 2)
                th = np.arctan2(x - cxn, y - Ly / 2)
                phi[..., n] = np.tanh((0.2 + 0.02 * np.cos(10 * th) - r) / s2e)
    psi = np.zeros_like(phi, order="F")
    for n in range(N):
        psi[..., n] = (-0.1 * phi[..., n] + 0.7) if conf.goal == "s" else (-0.35 * phi[..., n] + 0.45)
    return phi, psi


def generate_initial_condition(conf: Config, ops: Optional[Operators], state_type: int, seed: int = 1234):
    """generate_initial_condition of Init.jl:74-217 (2D) and Init_3d.jl:67-153 (3D): the
    physical fields of initial_fields_2d / initial_fields_3d and their spectra."""
    if conf.ndim == 3:
        phi, psi = initial_fields_3d(conf, state_type)
    else:
        phi, psi = initial_fields_2d(conf, state_type, seed)
    return state_from_physical(conf, ops, phi, psi)


def initial_fields_2d(conf: Config, state_type: int, seed: int = 1234):
    """(phi, psi) in Julia shape (Nx, Ny, N) of Init.jl:74-198; state_type 4 draws its
    mode amplitudes and phases from SplitMix64(seed) in the order of Init.jl:151-152."""
    Nx, Ny, N = conf.Nx, conf.Ny, conf.N
    Lx, Ly, eps = conf.Lx, conf.Ly, conf.epsilon
    x = np.linspace(0, Lx - conf.dx, Nx)
    y = np.linspace(0, Ly - conf.dy, Ny)
    X, Y = np.meshgrid(x, y, indexing="ij")                 # (Nx, Ny) like Init.jl:83-84
    phi = np.zeros((Nx, Ny, N), order="F")
    s2e = math.sqrt(2) * eps
    if N == 1:
        cx, cy = [Lx / 2], [Ly / 2]
    elif N == 2:
        if conf.goal not in ("g", "s"):
            raise ValueError(f"what's the goal? now it is {conf.goal} (Init.jl:91)")
        cx = [0.75, 1.25] if conf.goal == "g" else [0.65, 1.35]
        cy = [Ly / 2, Ly / 2]
    else:
        cx = [Lx / 4, Lx / 2, 3 * Lx / 4]
        cy = [Ly / 2] * 3
    # Init.jl:108-187: N = 1 raises for state_type > 6 and leaves phi = 0 for the types
    # without a shape (<= 0, 5, 6); N >= 2 raises for state_type <= 6, for 7 with goal s and
    # 8 with goal g, and leaves phi = 0 for state_type > 8
    if N == 1:
        if state_type > 6:
            raise ValueError(f"state_type {state_type} invalid for N = 1 (Init.jl:109)")
        cxn, cyn = cx[0], cy[0]
        if state_type == 1:
            R = 0.2 * Lx
            dist = np.sqrt((X - cxn) ** 2 / 2 + (Y - cyn) ** 2)
            phi[:, :, 0] = np.tanh((R - dist) / s2e)
        elif state_type == 2:
            Xc, Yc = X - cxn, Y - cyn
            radius, sm = 0.65 * Lx, 40.0
            ks, kc = math.sqrt(3) / 2, 0.5
            d1, d2, d3 = -Yc, ks * Xc + kc * Yc, -ks * Xc + kc * Yc
            mx = np.maximum(d1, np.maximum(d2, d3))
            d = mx + (1 / sm) * np.log(np.exp(sm * (d1 - mx)) + np.exp(sm * (d2 - mx)) + np.exp(sm * (d3 - mx)))
            phi[:, :, 0] = -np.tanh((d - radius / 3) / s2e)
        elif state_type == 3:
            R0, amp = (0.3, 0.01) if conf.goal == "s" else (0.18, 0.02)
            r = np.sqrt((X - cxn) ** 2 + (Y - cyn) ** 2)
            th = np.arctan2(X - cxn, Y - cyn)
            phi[:, :, 0] = np.tanh((R0 + amp * np.cos(10 * th) - r) / s2e)
        elif state_type == 4:
            R0, amp = (0.3, 0.15) if conf.goal == "s" else (0.2, 0.25)
            r = np.sqrt((X - cxn) ** 2 + (Y - cyn) ** 2)
            th = np.arctan2(X - cxn, Y - cyn)
            Rt = np.full(th.shape, R0)
            rng = SplitMix64(seed)
            for k in range(1, 11):
                an = (rng.uniform() * 2 - 1) * amp / k
                pn = rng.uniform() * 2 * math.pi
                Rt = Rt + an * R0 * np.cos(k * th + pn)
            phi[:, :, 0] = np.tanh((Rt - r) / s2e)
    else:
        if state_type <= 6:
            raise ValueError(f"state_type {state_type} invalid for N = {N} (Init.jl:157)")
        if state_type == 7 and conf.goal == "s":
            raise ValueError("state_type 7 needs goal g (Init.jl:163)")
        if state_type == 8 and conf.goal == "g":
            raise ValueError("state_type 8 needs goal s (Init.jl:174)")
        if N > 3:
            raise ValueError(f"N = {N}: Init.jl defines vesicle centres for N <= 3 only (Init.jl:86-100)")
        for n in range(N if state_type in (7, 8) else 0):
            R0, amp = (0.2, 0.02) if state_type == 7 else (0.3, 0.01)
            r = np.sqrt((X - cx[n]) ** 2 + (Y - cy[n]) ** 2)
            th = np.arctan2(X - cx[n], Y - cy[n])
            phi[:, :, n] = np.tanh((R0 + amp * np.cos(10 * th) - r) / s2e)
    psi = np.zeros_like(phi, order="F")
    for n in range(N):
        psi[:, :, n] = (-0.1 * phi[:, :, n] + 0.7) if conf.goal == "s" else (-0.35 * phi[:, :, n] + 0.45)
    return phi, psi


def state_from_physical(conf: Config, ops: Optional[Operators], phi: np.ndarray, psi: np.ndarray) -> FieldState:
    """A FieldState whose spectra are rfft of the physical fields (Init.jl:200-216);
    u = p = mu = nu = 0 and Q = R1 = R2 = R3 = 1.  ``phi``/``psi`` in Julia shape."""
    st = FieldState.zeros(conf)
    st.phi[...] = phi
    st.psi[...] = psi
    if ops is None:
        raise ValueError("state_from_physical needs the device operators for the transforms")
    st.phi_hat[...] = ops.rfft_accurate(st.phi)
    st.psi_hat[...] = ops.rfft_accurate(st.psi)
    st.Q = st.R1 = st.R2 = st.R3 = 1.0
    return st


def julia_shape(conf: Config, a: np.ndarray) -> np.ndarray:
    """Memory-order array (N, [Nz,] Ny, Nx) -> Julia-shaped Fortran view (Nx, Ny, [Nz,] N)."""
    return a.T


# ---------------------------------------------------------------------------
# run_simulation (src/simulation.jl:2-155) without plots
# ---------------------------------------------------------------------------
def time_nodes(dt: float, T: float) -> np.ndarray:
    """Two-phase schedule of simulation.jl:15-19: 0:1e-7:1e-5, then 1e-5:dt:T (first dropped)."""
    t1 = np.arange(0, 101) * 1e-7
    n2 = int(math.floor((T - 1e-5) / dt + 1e-9)) + 1
    t2 = 1e-5 + np.arange(n2) * dt
    return np.concatenate([t1, t2[1:]])


# ---------------------------------------------------------------------------
# Snapshots (simulation.jl:112-143) and checkpoints
# ---------------------------------------------------------------------------
SNAPSHOT_KEYS = ("phi", "psi", "mu", "nu", "p", "u", "R1", "R2", "R3", "Q")


def save_steps(Dt: np.ndarray, save_interval: float) -> List[int]:
    """1-based loop indices n after which run_simulation writes a snapshot: the rule of
    simulation.jl:113-114, curr = floor((Dt[n] + 1e-12) / save_interval), save when it
    exceeds the last saved index (initially -1).  Dt[n] is the time at the START of step
    n (Julia 1-based), so the snapshot holds the state after step n under that time."""
    out, last = [], -1
    for n in range(1, len(Dt)):
        cur = math.floor((float(Dt[n - 1]) + 1e-12) / save_interval)
        if cur > last:
            out.append(n)
            last = cur
    return out


def snapshot_name(conf: Config, t: float) -> str:
    """state_2d_t<t>.mat / state_3d_t<t>.mat with t printed as %.2e (simulation.jl:115,134-137)."""
    return f"state_{conf.ndim}d_t{t:.2e}.mat"


def write_snapshot(path: os.PathLike, st: FieldState, version: str = "7.3"):
    """The keys matwrite stores (simulation.jl:118-129), arrays in the Julia shapes, scalars
    as 1 x 1.  ``version`` "7.3" (the default of MAT.jl's matwrite, which the reference
    calls): the HDF5-based format, written by `matfile.write_mat73`; "5": MAT version 5
    through scipy.io (compressed), which MAT.jl's matread reads as well."""
    data = {k: getattr(st, k) for k in SNAPSHOT_KEYS}
    if version == "7.3":
        from .matfile import write_mat73
        write_mat73(path, data)
    elif version == "5":
        import scipy.io
        scipy.io.savemat(str(path), data, do_compression=True, oned_as="column")
    else:
        raise ValueError(f"write_snapshot: MAT version {version!r} is not written (7.3 or 5)")


def read_snapshot(path: os.PathLike) -> dict:
    """The snapshot keys of a MAT file written by write_snapshot (either version)."""
    from .matfile import is_mat73, read_mat73
    if is_mat73(path):
        d = read_mat73(path)
    else:
        import scipy.io
        d = scipy.io.loadmat(str(path))
    return {k: (float(d[k].ravel()[0]) if k in ("R1", "R2", "R3", "Q") else d[k]) for k in SNAPSHOT_KEYS}


def write_checkpoint(path: os.PathLike, present: FieldState, old: FieldState, n_next: int, t_next: float):
    """Everything a resume needs (no reference counterpart: the reference's snapshots drop
    `old`, A0 and the spectra): both states with spectra, Q/R, A0, dt, and the next loop
    index and time.  numpy .npz, nothing executable inside."""
    data = {"n_next": np.array(n_next), "t_next": np.array(t_next)}
    for tag, st in (("present", present), ("old", old)):
        for k in FieldState.ARRAYS:
            data[f"{tag}_{k}"] = np.asarray(getattr(st, k))
        data[f"{tag}_scal"] = np.array([st.Q, st.R1, st.R2, st.R3, st.dt])
        data[f"{tag}_A0"] = np.asarray(st.A0)
    np.savez_compressed(str(path), **data)


def read_checkpoint(path: os.PathLike):
    d = np.load(str(path), allow_pickle=False)
    out = []
    for tag in ("present", "old"):
        kw = {k: np.asfortranarray(d[f"{tag}_{k}"]) for k in FieldState.ARRAYS}
        Q, R1, R2, R3, dt = (float(v) for v in d[f"{tag}_scal"])
        out.append(FieldState(**kw, Q=Q, R1=R1, R2=R2, R3=R3, A0=np.array(d[f"{tag}_A0"]), dt=dt))
    return out[0], out[1], int(d["n_next"]), float(d["t_next"])


def run_simulation(dt_val: float, T_val: float, state_type: int, save_path: Optional[str] = None,
                   save_interval: float = 0.001, conf: Optional[Config] = None,
                   state: Optional[FieldState] = None, t_nodes: Optional[np.ndarray] = None,
                   precision: int = SAV_PREC_FP64, checkpoint: bool = False, plots: bool = True,
                   graph: int = GRAPH_AUTO):
    """Returns (present, energy_history, area_ratio_history, Dt) like simulation.jl:154.

    With ``conf``/``state``/``t_nodes`` left out it runs the reference's own set-up
    (set_para_base, generate_initial_condition, two-phase Dt schedule).  With
    ``save_path`` it writes the reference's snapshots (``save_path/MAT/case3/``, the
    schedule of simulation.jl:113-114), with ``plots`` the 2D phi image of each snapshot
    (``save_path/phi_t<t>.png``, simulation.jl:130-135) and, with ``checkpoint``, a resumable
    checkpoint next to each snapshot.  ``graph``: step replay mode (Operators.set_graph; by
    default chosen by measurement)."""
    conf = conf or set_para_base(dt_val, T_val)
    ops = build_operators(conf, precision=precision)
    ops.set_graph(graph)
    Dt = time_nodes(conf.dt, T_val) if t_nodes is None else np.asarray(t_nodes, dtype=np.float64)
    present = state if state is not None else generate_initial_condition(conf, ops, state_type)
    ops.set_state(SAV_PRESENT, present)
    ops.init_sav()
    E, A = [], []
    if save_path is None:
        rc, e, a, kit = ops.run(Dt, 1)
        E, A = list(e), list(a)
    else:
        mat_folder = Path(save_path) / "MAT" / "case3"
        mat_folder.mkdir(parents=True, exist_ok=True)
        saves = set(save_steps(Dt, save_interval))
        n0 = 1
        for n in sorted(saves | {len(Dt) - 1}):
            # steps n0 .. n (loop indices), time nodes Dt[n0-1] .. Dt[n]
            rc, e, a, kit = ops.run(Dt[n0 - 1: n + 1], n0)
            E += list(e)
            A += list(a)
            n0 = n + 1
            if n in saves:
                st = ops.get_state(SAV_PRESENT)
                write_snapshot(mat_folder / snapshot_name(conf, float(Dt[n - 1])), st)
                if plots and conf.ndim == 2:
                    from .plots import plot_field, phi_sum, _plt
                    t_str = f"{float(Dt[n - 1]):.2e}"
                    plot_field(phi_sum(st.phi, conf.N), conf, title=f"phi  (t = {t_str})",
                               filename=str(Path(save_path) / f"phi_t{t_str}.png"))
                    _plt().close("all")
                if checkpoint:
                    write_checkpoint(mat_folder / (snapshot_name(conf, float(Dt[n - 1]))[:-4] + ".ckpt.npz"),
                                     st, ops.get_state(SAV_OLD), n + 1, float(Dt[n]))
    present = ops.get_state(SAV_PRESENT)
    ops.close()
    return present, E, A, Dt


def resume_simulation(checkpoint_path: os.PathLike, conf: Config, t_nodes: np.ndarray,
                      precision: int = SAV_PREC_FP64):
    """Continues a run from write_checkpoint's file over the remaining time nodes
    (t_nodes[0] must be the checkpoint's time); returns (present, E, A)."""
    present, old, n_next, t_next = read_checkpoint(checkpoint_path)
    t = np.asarray(t_nodes, dtype=np.float64)
    if abs(t[0] - t_next) > 1e-12 * max(1.0, abs(t_next)):
        raise ValueError(f"t_nodes[0] = {t[0]} is not the checkpoint time {t_next}")
    ops = build_operators(conf, precision=precision)
    ops.set_state(SAV_PRESENT, present)
    ops.set_state(SAV_OLD, old)
    rc, E, A, kit = ops.run(t, n_next)
    out = ops.get_state(SAV_PRESENT)
    ops.close()
    return out, list(E), list(A)
