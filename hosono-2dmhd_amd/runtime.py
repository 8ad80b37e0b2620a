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
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# HMHD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = pathlib.Path(os.environ.get("HMHD_LIB", HERE / "libhmhd.so"))

BC_PERIODIC, BC_OUTFLOW = 0, 1
RIEMANN_HLLC = 1
# "tiles" holds counts: (MOOD iterations, active tiles they visited)
PROFILE_FAMILIES = ("face_full", "face_flux", "update", "recon", "flags", "cfl", "tiles")
RUN_ACTIVE, RUN_DONE, RUN_ABORTED = 0, 1, 2
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_ABORTED, REQ_FAILED = 0, 1, 2, 3, 4
ERR_TIMEOUT = 4

# every entry point include/hmhd.h declares
EXPORTS = (
    "hmhd_create", "hmhd_destroy", "hmhd_last_error", "hmhd_dims", "hmhd_gp_compute", "hmhd_get_gp",
    "hmhd_set_gp",
    "hmhd_upload_state", "hmhd_cfl_dt", "hmhd_advance", "hmhd_download", "hmhd_stats",
    "hmhd_record_stage_schemes", "hmhd_stage_schemes", "hmhd_profile", "hmhd_profile_read",
    "hmhd_synchronize", "hmhd_cfl_dt_noalloc", "hmhd_face_tuning", "hmhd_face_tuning_read",
    "hmhd_face_tuning_clear",
    "hmhd_batch_create", "hmhd_batch_destroy", "hmhd_batch_last_error", "hmhd_batch_add", "hmhd_batch_run",
    "hmhd_batch_status", "hmhd_batch_dims", "hmhd_batch_download", "hmhd_batch_profile",
    "hmhd_batch_profile_read",
    "hmhd_solve_footprint", "hmhd_server_create", "hmhd_server_destroy", "hmhd_server_last_error",
    "hmhd_server_submit", "hmhd_server_poll", "hmhd_server_wait", "hmhd_server_stats_read",
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


class _Result(C.Structure):
    _fields_ = [("state", C.c_int32), ("steps", C.c_int32), ("t", C.c_double), ("zone_updates", C.c_int64),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double)]


class _FaceTune(C.Structure):
    _fields_ = [("R", C.c_int32), ("size_class", C.c_int32), ("form", C.c_int32), ("tile_x", C.c_int32),
                ("tile_y", C.c_int32), ("waves", C.c_int32), ("ms", C.c_double * 8)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("active", C.c_int32),
                ("queued", C.c_int32), ("peak_active", C.c_int32), ("reserved", C.c_int32),
                ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64), ("engine_steps", C.c_int64),
                ("zone_updates", C.c_int64), ("busy_s", C.c_double), ("elapsed_s", C.c_double)]


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
    """Load libhmhd.so; raises if it has not been built.  With `path`, the
    library at that path is loaded and becomes the one every later Solver,
    Batch and Server of this process uses."""
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
    lib.hmhd_cfl_dt_noalloc.argtypes = [C.POINTER(_Config), vp, vp, vp, vp]
    lib.hmhd_face_tuning.argtypes = [C.c_int32]
    lib.hmhd_face_tuning_read.argtypes = [C.POINTER(_FaceTune), C.c_int32, C.POINTER(C.c_int32)]
    lib.hmhd_face_tuning_clear.argtypes = []
    i64, i64p = C.c_int64, C.POINTER(C.c_int64)
    lib.hmhd_batch_create.argtypes = [C.POINTER(vp)]
    lib.hmhd_batch_destroy.argtypes = [vp]
    lib.hmhd_batch_destroy.restype = None
    lib.hmhd_batch_last_error.argtypes = [vp]
    lib.hmhd_batch_last_error.restype = C.c_char_p
    lib.hmhd_batch_add.argtypes = [vp, C.POINTER(_Config), dp, dp, C.c_double, C.c_int32, i64p]
    lib.hmhd_batch_run.argtypes = [vp, C.c_int32]
    lib.hmhd_batch_status.argtypes = [vp, i64, dp, ip, ip]
    lib.hmhd_batch_dims.argtypes = [vp, i64, ip, ip, ip]
    lib.hmhd_batch_download.argtypes = [vp, i64, dp, dp, ip]
    lib.hmhd_batch_profile.argtypes = [vp, C.c_int32]
    lib.hmhd_batch_profile_read.argtypes = [vp, dp, i64p, C.c_int32]
    lib.hmhd_solve_footprint.argtypes = [C.POINTER(_Config)]
    lib.hmhd_solve_footprint.restype = C.c_uint64
    lib.hmhd_server_create.argtypes = [C.c_uint64, C.c_int32, C.POINTER(vp)]
    lib.hmhd_server_destroy.argtypes = [vp]
    lib.hmhd_server_destroy.restype = None
    lib.hmhd_server_last_error.argtypes = [vp]
    lib.hmhd_server_last_error.restype = C.c_char_p
    lib.hmhd_server_submit.argtypes = [vp, C.POINTER(_Config), dp, dp, C.c_double, C.c_int32, i64p]
    lib.hmhd_server_poll.argtypes = [vp, i64, C.POINTER(_Result)]
    lib.hmhd_server_wait.argtypes = [vp, i64, C.c_double, C.POINTER(_Result), dp, dp, ip]
    lib.hmhd_server_stats_read.argtypes = [vp, C.POINTER(_ServerStats)]
    _LIB = lib
    return lib


def _family_mask(families) -> int:
    if families is True:
        return (1 << len(PROFILE_FAMILIES)) - 1
    if not families:
        return 0
    return sum(1 << PROFILE_FAMILIES.index(f) for f in families)


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

    def profile(self, families=True):
        """Time kernel families with HIP events: True for all, False for
        none, or an iterable of names from PROFILE_FAMILIES."""
        self._check(self.lib.hmhd_profile(self.h, _family_mask(families)))

    def profile_read(self, reset: bool = False) -> dict:
        ms = np.zeros(len(PROFILE_FAMILIES))
        n = np.zeros(len(PROFILE_FAMILIES), dtype=np.int64)
        self._check(self.lib.hmhd_profile_read(self.h, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_int64)),
                                               int(reset)))
        return {f: (float(ms[k]), int(n[k])) for k, f in enumerate(PROFILE_FAMILIES)}

    def synchronize(self):
        self._check(self.lib.hmhd_synchronize(self.h))


def cfl_dt_noalloc(cfg: Config, U_ptr: int, V_ptr: int | None, dt_ptr: int, stream: int | None = None) -> None:
    """hmhd_cfl_dt_noalloc: enqueue the CFL dt of a state in device memory
    (U(8,mx,my) at U_ptr, V(10,mx,my) at V_ptr or None) into the device
    double at dt_ptr, on `stream` (a hipStream_t as an int; None: the null
    stream), allocating nothing."""
    lib = load_library()
    c = cfg.to_c()
    rc = lib.hmhd_cfl_dt_noalloc(C.byref(c), C.c_void_p(U_ptr), C.c_void_p(V_ptr) if V_ptr else None,
                                 C.c_void_p(dt_ptr), C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise HmhdError(rc, lib.hmhd_last_error(None).decode())


FACE_FORMS = ((16, 16, 2), (32, 8, 2), (32, 16, 2), (16, 16, 1), (32, 8, 1))


def face_tuning(mode: int | str = "tune") -> None:
    """The first-pass face kernel's form: "tune" (measure per GP radius and
    launch size class, keep the fastest; the default), "default" (16 x 16,
    two waves) or a form index into FACE_FORMS."""
    m = 1 if mode == "tune" else 0 if mode == "default" else 2 + int(mode)
    lib = load_library()
    rc = lib.hmhd_face_tuning(m)
    if rc != 0:
        raise HmhdError(rc, lib.hmhd_last_error(None).decode())


def face_tuning_table() -> list[dict]:
    """The kept forms: R, size class, form, its tile and waves, and the
    measured milliseconds per form."""
    lib = load_library()
    n = C.c_int32()
    buf = (_FaceTune * 64)()
    lib.hmhd_face_tuning_read(buf, 64, C.byref(n))
    out = []
    for e in buf[:min(n.value, 64)]:
        out.append({"R": e.R, "size_class": e.size_class, "form": e.form, "tile": [e.tile_x, e.tile_y],
                    "waves": e.waves,
                    "ms_per_form": {f"{FACE_FORMS[v][0]}x{FACE_FORMS[v][1]}/w{FACE_FORMS[v][2]}": e.ms[v]
                                    for v in range(len(FACE_FORMS)) if e.ms[v] > 0}})
    return out


def face_tuning_clear() -> None:
    load_library().hmhd_face_tuning_clear()


def _padded(cfg: Config):
    ngc = cfg.R + 2
    return cfg.nx + 2 * ngc, cfg.ny + 2 * ngc


def _state_args(cfg: Config, U, V):
    mx, my = _padded(cfg)
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.shape != (my, mx, 8):
        raise ValueError(f"U must have shape {(my, mx, 8)}, got {U.shape}")
    if V is not None:
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.shape != (my, mx, 10):
            raise ValueError(f"V must have shape {(my, mx, 10)}, got {V.shape}")
    return U, V


def footprint(cfg: Config) -> int:
    """Device bytes one run of this configuration holds."""
    c = cfg.to_c()
    return int(load_library().hmhd_solve_footprint(C.byref(c)))


class Batch:
    """Many independent runs of the driver loop (src/mhd_driver.f90:76-95)
    advanced together on one device; each keeps its own dt, t and end."""

    def __init__(self):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.hmhd_batch_create(C.byref(h))
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_batch_last_error(None).decode())
        self.h = h
        self.cfgs = {}

    def _check(self, rc):
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.hmhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, cfg: Config, U, V=None, tmax: float = 1.0, nstepmax: int = 0) -> int:
        U, V = _state_args(cfg, U, V)
        c = cfg.to_c()
        i = C.c_int64()
        self._check(self.lib.hmhd_batch_add(self.h, C.byref(c), _dp(U), _dp(V) if V is not None else None,
                                            float(tmax), int(nstepmax), C.byref(i)))
        self.cfgs[i.value] = cfg
        return i.value

    def run(self, max_steps: int = 0) -> None:
        """Advance every unfinished run by up to max_steps steps (0: to the end)."""
        self._check(self.lib.hmhd_batch_run(self.h, int(max_steps)))

    def status(self, rid: int):
        """(t, steps, state) with state one of RUN_ACTIVE, RUN_DONE, RUN_ABORTED."""
        t, n, st = C.c_double(), C.c_int32(), C.c_int32()
        self._check(self.lib.hmhd_batch_status(self.h, rid, C.byref(t), C.byref(n), C.byref(st)))
        return t.value, n.value, st.value

    def download(self, rid: int):
        mx, my = _padded(self.cfgs[rid])
        U = np.zeros((my, mx, 8))
        V = np.zeros((my, mx, 10))
        S = np.zeros((my, mx), dtype=np.int32)
        self._check(self.lib.hmhd_batch_download(self.h, rid, _dp(U), _dp(V), _ip(S)))
        return U, V, S

    def profile(self, families=True):
        self._check(self.lib.hmhd_batch_profile(self.h, _family_mask(families)))

    def profile_read(self, reset: bool = False) -> dict:
        ms = np.zeros(len(PROFILE_FAMILIES))
        n = np.zeros(len(PROFILE_FAMILIES), dtype=np.int64)
        self._check(self.lib.hmhd_batch_profile_read(self.h, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_int64)),
                                                     int(reset)))
        return {f: (float(ms[k]), int(n[k])) for k, f in enumerate(PROFILE_FAMILIES)}


@dataclasses.dataclass
class Result:
    state: int
    steps: int
    t: float
    zone_updates: int
    submit_s: float
    admit_s: float
    done_s: float
    U: np.ndarray | None = None
    V: np.ndarray | None = None
    scheme: np.ndarray | None = None

    @property
    def latency_s(self) -> float:
        return self.done_s - self.submit_s


class Server:
    """Runs admitted by HBM budget and stepped together by a worker thread
    in the library; submit/poll/wait are safe from any thread."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.hmhd_server_create(int(hbm_budget_bytes), int(max_active), C.byref(h))
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_server_last_error(None).decode())
        self.h = h
        self.cfgs = {}

    def _check(self, rc):
        if rc != 0:
            raise HmhdError(rc, self.lib.hmhd_server_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.hmhd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, cfg: Config, U, V=None, tmax: float = 1.0, nstepmax: int = 0) -> int:
        U, V = _state_args(cfg, U, V)
        c = cfg.to_c()
        r = C.c_int64()
        self._check(self.lib.hmhd_server_submit(self.h, C.byref(c), _dp(U), _dp(V) if V is not None else None,
                                                float(tmax), int(nstepmax), C.byref(r)))
        self.cfgs[r.value] = cfg
        return r.value

    def poll(self, req: int) -> Result:
        res = _Result()
        self._check(self.lib.hmhd_server_poll(self.h, req, C.byref(res)))
        return Result(res.state, res.steps, res.t, res.zone_updates, res.submit_s, res.admit_s, res.done_s)

    def wait(self, req: int, timeout_s: float = 0.0, arrays: bool = True) -> Result:
        cfg = self.cfgs[req]
        mx, my = _padded(cfg)
        U = np.zeros((my, mx, 8)) if arrays else None
        V = np.zeros((my, mx, 10)) if arrays else None
        S = np.zeros((my, mx), dtype=np.int32) if arrays else None
        res = _Result()
        self._check(self.lib.hmhd_server_wait(self.h, req, float(timeout_s), C.byref(res),
                                              _dp(U) if arrays else None, _dp(V) if arrays else None,
                                              _ip(S) if arrays else None))
        del self.cfgs[req]
        return Result(res.state, res.steps, res.t, res.zone_updates, res.submit_s, res.admit_s, res.done_s,
                      U, V, S)

    def stats(self) -> dict:
        st = _ServerStats()
        self._check(self.lib.hmhd_server_stats_read(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_ if f != "reserved"}


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
