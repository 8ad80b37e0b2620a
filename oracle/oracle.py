"""ctypes front end of the test oracle.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Two CPU implementations of the hosono-2DMHD step sit behind it:

* ``Oracle``  -- liborc.so, the plain-C restatement in hmhd_oracle.c;
* ``RefHarness`` -- oracle/_ref/libref_{exact,fast}.so, the reference's own
  Fortran modules compiled by amdflang (oracle/Makefile) and driven through
  ref_harness.f90.  The reference keeps its state in module globals, so only
  one RefHarness may be live per library.

Both take and return the reference's padded Fortran layout, as numpy arrays
of shape (my, mx, 8) for U and (my, mx, 10) for V.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import pathlib
import subprocess
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
BC_PERIODIC, BC_OUTFLOW = 0, 1


@dataclasses.dataclass
class Config:
    """The par-file quantities the step reads (src/simulation.f90:43-58,
    src/grid.f90:58-63, src/GP.f90:58-69)."""

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

    def par_text(self, tmax: float = 1.0) -> str:
        """A .par file with every key the reference reads."""
        f = lambda v: repr(float(v))
        lines = [
            "sim_dataFileBaseName = 'data/oracle'",
            f"sim_gamma = {f(self.gamma)}",
            f"sim_smallPressure = {f(self.small_p)}",
            "sim_smallEnergy = 1.0e-12",
            f"sim_smallDensity = {f(self.small_rho)}",
            f"sim_smallB = {f(self.small_B)}",
            "sim_riemannSolver = 'hllc'",
            f"sim_cfl = {f(self.cfl)}",
            f"sim_tmax = {f(tmax)}",
            "sim_nstepmax = 1000000",
            f"sim_forceHydro = {'.true.' if self.force_hydro else '.false.'}",
            "sim_outputFreqTime = 0.0",
            "sim_outputFreqStep = 0",
            "sim_outputHdf5 = .false.",
            "sim_outputAscii = .false.",
            f"grid_Nx = {self.nx}",
            f"grid_Ny = {self.ny}",
            f"grid_xBeg = {f(self.xbeg)}",
            f"grid_xEnd = {f(self.xend)}",
            f"grid_yBeg = {f(self.ybeg)}",
            f"grid_yEnd = {f(self.yend)}",
            f"GP_maxRadius = {self.R}",
            f"GP_ellOverDelta = {f(self.ell_over_delta)}",
            f"sim_BC = '{'periodic' if self.bc == BC_PERIODIC else 'outflow'}'",
        ]
        return "\n".join(lines) + "\n"


class _OrcCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("xbeg", C.c_double), ("xend", C.c_double),
                ("ybeg", C.c_double), ("yend", C.c_double), ("R", C.c_int),
                ("ell_over_delta", C.c_double), ("gamma", C.c_double), ("small_p", C.c_double),
                ("small_rho", C.c_double), ("small_B", C.c_double), ("force_hydro", C.c_int),
                ("bc", C.c_int), ("cfl", C.c_double)]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def build(ref: bool = False) -> None:
    """Build liborc.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref and pathlib.Path("/root/reference/src").is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


_LIBORC = None


def _liborc():
    global _LIBORC
    if _LIBORC is None:
        path = HERE / "liborc.so"
        if not path.exists():
            build()
        lib = C.CDLL(str(path))
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [C.POINTER(_OrcCfg)]
        lib.orc_destroy.argtypes = [vp]
        lib.orc_dims.argtypes = [vp, ip, ip, ip]
        lib.orc_gp_tables.argtypes = [vp, dp, dp, ip]
        lib.orc_set_state.argtypes = [vp, dp, dp]
        lib.orc_get_state.argtypes = [vp, dp, dp]
        lib.orc_cfl_dt.restype = C.c_double
        lib.orc_cfl_dt.argtypes = [vp]
        lib.orc_advance.argtypes = [vp, C.c_double]
        lib.orc_get_stage_schemes.argtypes = [vp, ip]
        lib.orc_get_stats.argtypes = [vp, ip, ip]
        lib.orc_forward_euler.argtypes = [vp, C.c_double, dp, dp, ip]
        lib.orc_reconstruct_all.argtypes = [vp, dp, C.c_int, dp]
        lib.orc_bc_fill.argtypes = [vp, dp, C.c_int]
        lib.orc_hllc.argtypes = [vp, dp, dp, C.c_int, dp]
        lib.orc_cons2prim.argtypes = [vp, dp, dp]
        lib.orc_prim2cons.argtypes = [vp, dp, dp]
        lib.orc_cons2flux.argtypes = [vp, dp, C.c_int, dp]
        lib.orc_cons2pres.restype = C.c_double
        lib.orc_cons2pres.argtypes = [vp, dp]
        lib.orc_eigen_cfs.restype = C.c_double
        lib.orc_eigen_cfs.argtypes = [C.c_double] * 6
        _LIBORC = lib
    return _LIBORC


class Oracle:
    """The C restatement (hmhd_oracle.c)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.lib = _liborc()
        c = _OrcCfg(cfg.nx, cfg.ny, cfg.xbeg, cfg.xend, cfg.ybeg, cfg.yend, cfg.R,
                    cfg.ell_over_delta, cfg.gamma, cfg.small_p, cfg.small_rho, cfg.small_B,
                    int(cfg.force_hydro), cfg.bc, cfg.cfl)
        self.h = self.lib.orc_create(C.byref(c))
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        ngc, mx, my = C.c_int(), C.c_int(), C.c_int()
        self.lib.orc_dims(self.h, C.byref(ngc), C.byref(mx), C.byref(my))
        self.ngc, self.mx, self.my = ngc.value, mx.value, my.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def gp_tables(self):
        R = self.cfg.R
        ns, nq = 1 + 2 * R * (R + 1), R + 1
        pv = np.zeros((R, 4, nq, ns))
        qw = np.zeros((R, nq))
        st = np.zeros((R, ns, 2), dtype=np.int32)
        self.lib.orc_gp_tables(self.h, _dp(pv), _dp(qw), _ip(st))
        return pv, qw, st

    def set_state(self, U, V=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.my, self.mx, 8)
        if V is not None:
            V = np.ascontiguousarray(V, dtype=np.float64)
            assert V.shape == (self.my, self.mx, 10)
            self.lib.orc_set_state(self.h, _dp(U), _dp(V))
        else:
            self.lib.orc_set_state(self.h, _dp(U), None)

    def get_state(self):
        U = np.zeros((self.my, self.mx, 8))
        V = np.zeros((self.my, self.mx, 10))
        self.lib.orc_get_state(self.h, _dp(U), _dp(V))
        return U, V

    def cfl_dt(self) -> float:
        return self.lib.orc_cfl_dt(self.h)

    def advance(self, dt: float) -> None:
        self.lib.orc_advance(self.h, dt)

    def stage_schemes(self):
        s = np.zeros((3, self.my, self.mx), dtype=np.int32)
        self.lib.orc_get_stage_schemes(self.h, _ip(s))
        return s

    def stats(self):
        p = np.zeros(3, dtype=np.int32)
        q = np.zeros(3, dtype=np.int32)
        self.lib.orc_get_stats(self.h, _ip(p), _ip(q))
        return p, q

    def reconstruct_all(self, U, radius):
        U = np.ascontiguousarray(U, dtype=np.float64)
        fv = np.zeros((self.my, self.mx, 4, 4, 8))
        self.lib.orc_reconstruct_all(self.h, _dp(U), radius, _dp(fv))
        return fv  # [j, i, face, q, var]

    def hllc(self, uL, uR, d):
        out = np.zeros(8)
        self.lib.orc_hllc(self.h, _dp(np.ascontiguousarray(uL, dtype=np.float64)),
                          _dp(np.ascontiguousarray(uR, dtype=np.float64)), d, _dp(out))
        return out

    def cons2prim(self, U):
        out = np.zeros(10)
        self.lib.orc_cons2prim(self.h, _dp(np.ascontiguousarray(U, dtype=np.float64)), _dp(out))
        return out

    def prim2cons(self, V):
        out = np.zeros(8)
        self.lib.orc_prim2cons(self.h, _dp(np.ascontiguousarray(V, dtype=np.float64)), _dp(out))
        return out

    def cons2flux(self, U, d):
        out = np.zeros(8)
        self.lib.orc_cons2flux(self.h, _dp(np.ascontiguousarray(U, dtype=np.float64)), d, _dp(out))
        return out

    def cons2pres(self, U):
        return self.lib.orc_cons2pres(self.h, _dp(np.ascontiguousarray(U, dtype=np.float64)))

    def eigen_cfs(self, gam, rho, p, BB, BN, pm):
        return self.lib.orc_eigen_cfs(gam, rho, p, BB, BN, pm)

    def run(self, nsteps: int, tmax: float) -> tuple[float, int]:
        """Driver loop of src/mhd_driver.f90:76-95 (dt clamp to tmax)."""
        t, n = 0.0, 0
        while t < tmax and n < nsteps:
            dt = self.cfl_dt()
            if t + dt > tmax:
                dt = tmax - t
            self.advance(dt)
            t += dt
            n += 1
        return t, n


def ref_available(variant: str = "exact") -> bool:
    return (HERE / "_ref" / f"libref_{variant}.so").exists()


class RefHarness:
    """The reference's own Fortran step (oracle/_ref)."""

    _libs: dict = {}

    def __init__(self, cfg: Config | None, variant: str = "exact", par_path=None):
        """Set up the reference's grid from cfg, or, with par_path, from that
        .par file as the reference's driver reads it (cfg may be None)."""
        self.cfg = cfg
        path = HERE / "_ref" / f"libref_{variant}.so"
        if variant not in RefHarness._libs:
            lib = C.CDLL(str(path))
            vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
            lib.ref_init.argtypes = [C.c_char_p]
            lib.ref_dims.argtypes = [ip] * 6
            lib.ref_gp_tables.argtypes = [dp, dp, ip]
            lib.ref_set_state.argtypes = [dp, dp]
            lib.ref_get_state.argtypes = [dp, dp, ip]
            lib.ref_cfl_dt.restype = C.c_double
            lib.ref_advance.argtypes = [C.c_double]
            lib.ref_run.restype = C.c_double
            lib.ref_run.argtypes = [C.c_int, C.c_double]
            lib.ref_reconstruct_all.argtypes = [C.c_int, dp]
            lib.ref_step_masks.argtypes = [C.c_double, dp, ip, ip]
            lib.ref_flush.argtypes = []
            lib.ref_initial_condition.argtypes = [C.c_char_p]
            RefHarness._libs[variant] = lib
        self.lib = RefHarness._libs[variant]
        self._tmp = tempfile.TemporaryDirectory()
        if par_path is None:
            par = pathlib.Path(self._tmp.name) / "run.par"
            par.write_text(cfg.par_text())
        else:
            par = pathlib.Path(par_path)
        self.par = par
        # the reference prints banners on stdout through the Fortran runtime
        self.lib.ref_init(str(par).encode())
        d = [C.c_int() for _ in range(6)]
        self.lib.ref_dims(*[C.byref(x) for x in d])
        self.ngc, self.mx, self.my, self.nsmax, self.nqmax, self.rmax = (x.value for x in d)

    def flush(self) -> None:
        self.lib.ref_flush()

    def initial_condition(self):
        """src/initialCondition.f90:23-84 from this harness's .par file;
        returns (U, V) with guard cells, as get_state."""
        self.lib.ref_initial_condition(str(self.par).encode())
        U, V, _ = self.get_state()
        return U, V

    def gp_tables(self):
        pv = np.zeros((self.rmax, 4, self.nqmax, self.nsmax))
        qw = np.zeros((self.rmax, self.nqmax))
        st = np.zeros((self.rmax, self.nsmax, 2), dtype=np.int32)
        self.lib.ref_gp_tables(_dp(pv), _dp(qw), _ip(st))
        return pv, qw, st

    def set_state(self, U, V):
        self.lib.ref_set_state(_dp(np.ascontiguousarray(U, dtype=np.float64)),
                               _dp(np.ascontiguousarray(V, dtype=np.float64)))

    def get_state(self):
        U = np.zeros((self.my, self.mx, 8))
        V = np.zeros((self.my, self.mx, 10))
        s = np.zeros((self.my, self.mx), dtype=np.int32)
        self.lib.ref_get_state(_dp(U), _dp(V), _ip(s))
        return U, V, s

    def cfl_dt(self) -> float:
        return self.lib.ref_cfl_dt()

    def advance(self, dt: float) -> None:
        self.lib.ref_advance(dt)

    def run(self, nsteps: int, tmax: float) -> float:
        return self.lib.ref_run(nsteps, tmax)

    def reconstruct_all(self, radius: int):
        fv = np.zeros((self.my, self.mx, 4, self.nqmax, 8))
        self.lib.ref_reconstruct_all(radius, _dp(fv))
        return fv

    def step_masks(self, dt: float):
        U = np.zeros((self.my, self.mx, 8))
        s = np.zeros((3, self.my, self.mx), dtype=np.int32)
        p = np.zeros(3, dtype=np.int32)
        self.lib.ref_step_masks(dt, _dp(U), _ip(s), _ip(p))
        return U, s, p
