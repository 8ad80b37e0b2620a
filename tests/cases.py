"""Workloads used by the parity tests: (name, oracle Config, grid, U, V).

Every case is a solve the reference can run through its own step
(src/timeStep.f90) with initial data set from outside, as the tests do.
"""
import math

import hosono_2dmhd_amd.ics as ics
import oracle as O


def make(name, n=None, R=None):
    """Build one named case; n / R override the default size and radius."""
    if name == "ot":  # workload Orszag-Tang, periodic, R=2, l=12 dx
        n = n or 64
        R = R or 2
        g = ics.Grid(n, n, R=R)
        U, V = ics.orszag_tang(g)
        cfg = O.Config(n, n, R=R, ell_over_delta=12.0)
    elif name == "ot_rect":  # periodic, nx != ny, dx != dy
        R = R or 2
        g = ics.Grid(48, 64, 0.0, 1.0, 0.0, 1.0, R=R)
        U, V = ics.orszag_tang(g)
        cfg = O.Config(48, 64, 0.0, 1.0, 0.0, 1.0, R=R, ell_over_delta=12.0)
    elif name == "ot_ref":  # the reference's own OT IC and par (R=3, l=6 dx)
        n = n or 48
        R = R or 3
        gamma = 1.66666666666666666
        g = ics.Grid(n, n, R=R)
        U, V = ics.orszag_tang_reference(g, gamma)
        cfg = O.Config(n, n, R=R, ell_over_delta=6.0, gamma=gamma)
    elif name == "blast":  # MHD blast, outflow
        n = n or 64
        R = R or 2
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=R)
        U, V = ics.mhd_blast(g)
        cfg = O.Config(n, n, -0.5, 0.5, -0.5, 0.5, R=R, bc=O.BC_OUTFLOW)
    elif name == "blast_periodic":  # MHD blast centred on a corner of a periodic domain:
        n = n or 48                   # inner-loop rounds across the periodic guard bands
        R = R or 2
        g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=R)
        U, V = ics.mhd_blast(g)
        ics.bc_fill(U, g, ics.BC_PERIODIC)
        ics.bc_fill(V, g, ics.BC_PERIODIC)
        cfg = O.Config(n, n, 0.0, 1.0, 0.0, 1.0, R=R)
    elif name == "blast_rect":
        R = R or 2
        g = ics.Grid(64, 40, -0.5, 0.5, -0.3, 0.3, R=R)
        U, V = ics.mhd_blast(g)
        cfg = O.Config(64, 40, -0.5, 0.5, -0.3, 0.3, R=R, bc=O.BC_OUTFLOW)
    elif name == "sedov":  # workload Sedov with B, outflow
        n = n or 48
        R = R or 2
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=R)
        U, V = ics.sedov(g)
        cfg = O.Config(n, n, -0.5, 0.5, -0.5, 0.5, R=R, bc=O.BC_OUTFLOW)
    elif name == "explosion_ref":  # the reference's hydro explosion par
        n = n or 48
        R = R or 1
        g = ics.Grid(n, n, -1.0, 1.0, 0.0, 2.0, R=R)
        U, V = ics.explosion_reference(g, 1.4)
        cfg = O.Config(n, n, -1.0, 1.0, 0.0, 2.0, R=R, ell_over_delta=6.0, gamma=1.4,
                       force_hydro=True, bc=O.BC_OUTFLOW, cfl=0.5)
    elif name == "sedov_ref":  # the reference's hydro Sedov par
        n = n or 48
        R = R or 3
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=R)
        U, V = ics.sedov_reference(g, 1.4)
        cfg = O.Config(n, n, -0.5, 0.5, -0.5, 0.5, R=R, ell_over_delta=6.0, gamma=1.4,
                       force_hydro=True, bc=O.BC_OUTFLOW)
    else:
        raise KeyError(name)
    return cfg, g, U, V


def runtime_config(cfg):
    """The same configuration for the HIP library's front end."""
    import hosono_2dmhd_amd.runtime as rt
    return rt.Config(cfg.nx, cfg.ny, cfg.xbeg, cfg.xend, cfg.ybeg, cfg.yend, cfg.R,
                     cfg.ell_over_delta, cfg.gamma, cfg.small_p, cfg.small_rho, cfg.small_B,
                     cfg.force_hydro, cfg.bc, cfg.cfl)
