"""Python front end of libhmhd.so (the C ABI in include/hmhd.h).

Mirrors the reference's driver-facing interface: ``Solver.cfl_computedt``
replaces ``cfl_computedt`` (src/cfl.f90:14-52), ``Solver.advance_solution``
replaces ``timeStep_advanceSolution`` (src/timeStep.f90:22-43), and
``valid_time_step`` / ``run`` restate the driver loop of
src/mhd_driver.f90:76-144.  Arrays use the reference's padded Fortran
layout: numpy (my, mx, 8) for U(8, mx, my), (my, mx, 10) for V.

The library is built in-tree by ``make -C hosono-2dmhd_amd`` (or
``__graft_entry__.build()``); loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libhmhd.so"

BC_PERIODIC, BC_OUTFLOW = 0, 1
RIEMANN_HLLC = 1
PROFILE_FAMILIES = ("face_full", "face_flux", "update", "recon", "flags", "cfl")

# every entry point include/hmhd.h declares
EXPORTS = (
    "hmhd_create", "hmhd_destroy", "hmhd_last_error", "hmhd_dims", "hmhd_gp_compute", "hmhd_get_gp",
    "hmhd_set_gp",
    "hmhd_upload_state", "hmhd_cfl_dt", "hmhd_advance", "hmhd_download", "hmhd_stats",
    "hmhd_record_stage_schemes", "hmhd_stage_schemes", "hmhd_profile", "hmhd_profile_read",
    "hmhd_synchronize",
)


class HmhdError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"hmhd error {code}: {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("xbeg", C.c_double), ("xend", C.c_double),
                ("ybeg", C.c_double), ("yend", C.c_double), ("R", C.c_int32),
                ("ell_over_delta", C.c_double), ("gamma", C.c_double), ("small_p", C.c_double),
                ("small_rho", C.c_double), ("small_B", C.c_double), ("force_hydro", C.c_int32),
                ("bc", C.c_int32), ("cfl", C.c_double), ("riemann", C.c_int32)]


@dataclasses.dataclass
class Config:
    """The par-file quantities the step reads (see hmhd_config)."""

    nx: int
    ny: int
    xbeg: float = 0.0
    xend: float = 1.0
    ybeg: float = 0.0
    yend: float = 1.0
    R: int = 2
    ell_over_delta: float = 12.0
    gamma: float = 5.0 / 3.0
    small_p: float = 1e-12
    small_rho: float = 1e-12
    small_B: float = 1e-12
    force_hydro: bool = False
    bc: int = BC_PERIODIC
    cfl: float = 0.8

    def to_c(self) -> _Config:
        return _Config(self.nx, self.ny, self.xbeg, self.xend, self.ybeg, self.yend, self.R,
                       self.ell_over_delta, self.gamma, self.small_p, self.small_rho, self.small_B,
                       int(self.force_hydro), self.bc, self.cfl, RIEMANN_HLLC)


_LIB = None


def load_library(path: pathlib.Path | str | None = None):
    """Load libhmhd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    lib.hmhd_create.argtypes = [C.POINTER(_Config), C.POINTER(vp)]
    lib.hmhd_destroy.argtypes = [vp]
    lib.hmhd_destroy.restype = None
    lib.hmhd_last_error.argtypes = [vp]
    lib.hmhd_last_error.restype = C.c_char_p
    lib.hmhd_dims.argtypes = [vp, ip, ip, ip]
    lib.hmhd_gp_compute.argtypes = [C.POINTER(_Config), dp, dp, ip]
    lib.hmhd_get_gp.argtypes = [vp, dp, dp, ip]
    lib.hmhd_set_gp.argtypes = [vp, dp, dp]
    lib.hmhd_upload_state.argtypes = [vp, dp, dp]
    lib.hmhd_cfl_dt.argtypes = [vp, dp]
    lib.hmhd_advance.argtypes = [vp, C.c_double]
    lib.hmhd_download.argtypes = [vp, dp, dp, ip]
    lib.hmhd_stats.argtypes = [vp, ip, ip]
    lib.hmhd_record_stage_schemes.argtypes = [vp, C.c_int32]
    lib.hmhd_stage_schemes.argtypes = [vp, ip]
    lib.hmhd_profile.argtypes = [vp, C.c_int32]
    lib.hmhd_profile_read.argtypes = [vp, dp, C.POINTER(C.c_int64), C.c_int32]
    lib.hmhd_synchronize.argtypes = [vp]
    if path is None:
        _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def gp_compute(cfg: Config):
    """GP tables of a configuration computed on the host (no device):
    (predvec[R,4,nq,ns], quadw[R,nq], stenc[R,ns,2]), reference layout."""
    lib = load_library()
    R = cfg.R
    ns, nq = 1 + 2 * R * (R + 1), R + 1
    pv = np.zeros((R, 4, nq, ns))
    qw = np.zeros((R, nq))
    st = np.zeros((R, ns, 2), dtype=np.int32)
    c = cfg.to_c()
    rc = lib.hmhd_gp_compute(C.byref(c), _dp(pv), _dp(qw), _ip(st))
    if rc != 0:
        raise HmhdError(rc, lib.hmhd_last_error(None).decode())
    return pv, qw, st


class Solver:
    """One solve on the GPU: the grid, GP tables and device-resident state."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.lib = load_library()
        h = C.c_void_p()
        c = cfg.to_c()
        rc = self.lib.hmhd_create(C.byref(c), C.byref(h))
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_last_error(None).decode())
        self.h = h
        ngc, mx, my = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.hmhd_dims(self.h, C.byref(ngc), C.byref(mx), C.byref(my)))
        self.ngc, self.mx, self.my = ngc.value, mx.value, my.value

    def _check(self, rc):
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.hmhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- set-up ---------------------------------------------------------
    def gp_tables(self):
        """(predvec[R,4,nq,ns], quadw[R,nq], stenc[R,ns,2]) in the reference layout."""
        R = self.cfg.R
        ns, nq = 1 + 2 * R * (R + 1), R + 1
        pv = np.zeros((R, 4, nq, ns))
        qw = np.zeros((R, nq))
        st = np.zeros((R, ns, 2), dtype=np.int32)
        self._check(self.lib.hmhd_get_gp(self.h, _dp(pv), _dp(qw), _ip(st)))
        return pv, qw, st

    def set_gp_tables(self, pv, qw):
        pv = np.ascontiguousarray(pv, dtype=np.float64)
        qw = np.ascontiguousarray(qw, dtype=np.float64)
        self._check(self.lib.hmhd_set_gp(self.h, _dp(pv), _dp(qw)))

    def upload(self, U, V=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (self.my, self.mx, 8):
            raise ValueError(f"U must have shape {(self.my, self.mx, 8)}, got {U.shape}")
        if V is not None:
            V = np.ascontiguousarray(V, dtype=np.float64)
            if V.shape != (self.my, self.mx, 10):
                raise ValueError(f"V must have shape {(self.my, self.mx, 10)}, got {V.shape}")
        self._check(self.lib.hmhd_upload_state(self.h, _dp(U), _dp(V) if V is not None else None))

    # --- the hot path ---------------------------------------------------
    def cfl_computedt(self) -> float:
        dt = C.c_double()
        self._check(self.lib.hmhd_cfl_dt(self.h, C.byref(dt)))
        return dt.value

    def advance_solution(self, dt: float) -> None:
        self._check(self.lib.hmhd_advance(self.h, float(dt)))

    # --- output / observability ------------------------------------------
    def download(self, want_U=True, want_V=True, want_scheme=True):
        U = np.zeros((self.my, self.mx, 8)) if want_U else None
        V = np.zeros((self.my, self.mx, 10)) if want_V else None
        S = np.zeros((self.my, self.mx), dtype=np.int32) if want_scheme else None
        self._check(self.lib.hmhd_download(self.h, _dp(U) if U is not None else None,
                                           _dp(V) if V is not None else None,
                                           _ip(S) if S is not None else None))
        return U, V, S

    def stats(self):
        p = np.zeros(3, dtype=np.int32)
        q = np.zeros(3, dtype=np.int32)
        self._check(self.lib.hmhd_stats(self.h, _ip(p), _ip(q)))
        return p, q

    def record_stage_schemes(self, on: bool = True):
        self._check(self.lib.hmhd_record_stage_schemes(self.h, int(on)))

    def stage_schemes(self):
        s = np.zeros((3, self.my, self.mx), dtype=np.int32)
        self._check(self.lib.hmhd_stage_schemes(self.h, _ip(s)))
        return s

    def profile(self, on: bool = True):
        self._check(self.lib.hmhd_profile(self.h, int(on)))

    def profile_read(self, reset: bool = False) -> dict:
        ms = np.zeros(len(PROFILE_FAMILIES))
        n = np.zeros(len(PROFILE_FAMILIES), dtype=np.int64)
        self._check(self.lib.hmhd_profile_read(self.h, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_int64)),
                                               int(reset)))
        return {f: (float(ms[k]), int(n[k])) for k, f in enumerate(PROFILE_FAMILIES)}

    def synchronize(self):
        self._check(self.lib.hmhd_synchronize(self.h))


class TimeStepError(RuntimeError):
    """The driver's validTimeStep abort (src/mhd_driver.f90:117-144)."""


def valid_time_step(dt: float, t: float, tmax: float) -> float:
    """src/mhd_driver.f90:117-144: clamp to tmax; abort on a negative,
    NaN or infinite dt."""
    if t + dt > tmax:
        return tmax - t
    if dt < 0.0:
        raise TimeStepError("dt is negative. Aborting simulation.")
    if math.isnan(dt):
        raise TimeStepError("dt is NAN. Aborting simulation.")
    if not math.isfinite(dt):
        raise TimeStepError("dt is infinite. Aborting simulation.")
    return dt


def run(solver: Solver, tmax: float, nstepmax: int = 1_000_000, on_step=None) -> tuple[float, int]:
    """The main loop of src/mhd_driver.f90:64-95 without output: the first
    CFL call, then cfl -> validTimeStep -> advance until tmax or nstepmax."""
    t, n = 0.0, 0
    valid_time_step(solver.cfl_computedt(), t, tmax)
    while t < tmax and n < nstepmax:
        dt = valid_time_step(solver.cfl_computedt(), t, tmax)
        solver.advance_solution(dt)
        t += dt
        n += 1
        if on_step is not None:
            on_step(n, t, dt)
    return t, n
