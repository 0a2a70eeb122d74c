"""Workload definitions: the three reference scenarios and the throughput block.

Each builder returns a :class:`Scenario` holding the model parameters, the
reference coordinates and the setup sequence of the matching reference solver.
"""
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from .engine import Problem, gen_block_jitter, gen_kw, gen_plate, gen_plate_strips, params


@dataclass
class Scenario:
    name: str
    p: _lib.PdParams
    x: np.ndarray
    y: np.ndarray
    z: Optional[np.ndarray]
    n_int: int
    steps: int
    nxyz: tuple
    setup: Callable[[Problem], None] = field(repr=False, default=None)
    cyl: Optional[tuple] = None            # rad, mass, height, v, x, y, z
    v0: Optional[np.ndarray] = None         # initial velocity [dim][N]

    def make(self, device=0, precision=None, group=0):
        p = self.p
        if precision is not None:
            p.precision = precision
        prob = Problem(p, device)
        prob.set_positions(self.x, self.y, self.z, n_int=self.n_int)
        prob.build_neighbors()
        self.setup(prob)
        if self.cyl is not None:
            prob.cylinder_init(*self.cyl)
        if self.v0 is not None:
            prob.set("V", self.v0)
        return prob


def c0_plate_pmb(scale=1.0, steps=1000):
    """2D bond-based PMB plate with boundary strips pulled at +-20 m/s.

    Layout of Coord.cu:47-66 and loading of Solve_Plate_Crack_Grow.cu:8-27,128;
    PMB force of Force_Bond.cu:159-202.  50x50 mm at dx = 0.5 mm (scale 1),
    t = dx, E = 192 GPa, rho = 8000, nu = 1/3, s0 = 0.004, dt = 1e-8 s,
    10 mm central crack.
    """
    dx = 0.5e-3
    L = 0.05 * scale
    p, nxyz, _ = params(2, _lib.PD_MODEL_PMB, 1.0e-8, dx, thick=dx, length=L, height=L, emod=192e9,
                        pratio=1.0 / 3.0, dens=8000.0, s0=0.004)
    nx, ny, _ = nxyz
    p.n_top = p.n_bot = 3 * nx
    p.strip_velocity = 20.0
    x, y = gen_plate_strips(nx, ny, dx)
    clen = 0.01 * scale

    def setup(prob):
        prob.init_state()
        prob.set_crack_2d(clen, 0.0, 0.0, 0.0)

    return Scenario("c0_plate_pmb", p, x, y, None, nx * ny, steps, nxyz, setup)


def lps2d_plate(scale=1.0, steps=1000, precision=None):
    """2D state-based LPS plate with strips pulled at +-20 m/s (the LPS branch
    of Solve_Plate_Crack_Grow.cu:100-167, Force_State.cu:309-393): 6 x 6 mm at
    dx = 0.1 mm (scale 1), t = dx, E = 192 GPa, nu = 0.3, rho = 7850,
    s0 = 0.04472, dt = 1.3367e-8 s, 2 mm central crack."""
    dx = 1e-4
    L = 0.006 * scale
    p, nxyz, _ = params(2, _lib.PD_MODEL_LPS, 1.3367e-8, dx, thick=1e-4, length=L, height=L, emod=192e9,
                        pratio=0.3, dens=7850.0, s0=0.04472)
    if precision is not None:
        p.precision = precision
    nx, ny, _ = nxyz
    p.n_top = p.n_bot = 3 * nx
    p.strip_velocity = 20.0
    x, y = gen_plate_strips(nx, ny, dx)
    clen = 0.002 * scale

    def setup(prob):
        prob.init_state()
        prob.set_crack_2d(clen, 0.0, 0.0, 0.0)
        prob.lps_weights()

    return Scenario("lps2d_plate", p, x, y, None, nx * ny, steps, nxyz, setup)


def kw_s0(emod, pratio, delta, G0=22170.0):
    """Critical stretch of 3D PMB from the fracture energy: sqrt(5 G0 / (9 K delta))."""
    K = emod / (3.0 * (1.0 - 2.0 * pratio))
    return math.sqrt(5.0 * G0 / (9.0 * K * delta))


def c1_kalthoff_winkler(steps=1350, delta_factor=3.0, width=0.009):
    """3D PMB Kalthoff-Winkler impact (Solve_Kalthoff_Winkler.cu:268-441).

    200x100x9 mm at dx = 1 mm, grid (nx+1)x(ny+1)xnz with the two notches of
    Coord.cu:77-129, surface correction, notch bonds cut
    (Solve_Kalthoff_Winkler.cu:65-87), break region y >= -0.04, rigid cylinder
    impactor (R 25 mm, H 50 mm, M 1.57 kg, v -32 m/s), dt = 8.7e-8 s.
    """
    dx = 1.0e-3
    emod, nu = 190e9, 0.25
    delta = delta_factor * dx
    s0 = kw_s0(emod, nu, delta)
    p, nxyz, _ = params(3, _lib.PD_MODEL_PMB, 8.7e-8, dx, thick=dx, length=0.2, height=0.1, width=width,
                        emod=emod, pratio=nu, dens=8000.0, s0=s0, delta_factor=delta_factor)
    nx, ny, nz = nxyz[0] + 1, nxyz[1] + 1, nxyz[2]       # Solve_Kalthoff_Winkler.cu:282-284
    rad = 0.025
    x, y, z = gen_kw(nx, ny, nz, dx, rad)
    p.break_y_min = -0.04
    cyl_h = 0.05
    cyl = (rad, 1.57, cyl_h, -32.0, 0.0, 0.1 / 2.0 + cyl_h / 2.0 + 0.1 * dx, 0.0)

    def setup(prob):
        prob.surface_correct()
        prob.init_state()
        prob.set_notches_kw(rad)

    return Scenario("c1_kalthoff_winkler", p, x, y, z, len(x), steps, (nx, ny, nz), setup, cyl=cyl)


def c2_lps_block(nx=200, ny=200, nz=100, steps=200, precision=_lib.PD_PREC_MIXED, seed=0x5EED0002):
    """3D state-based LPS block: 0.1 x 0.1 x 0.05 m at dx = 0.5 mm (4M particles).

    Lattice with seeded +-1% dx jitter, E = 70 GPa, nu = 0.33, rho = 2700,
    s0 = 0.002, delta = 3.015 dx, K = E/(3(1-2nu)), v_z(0) = 1e3 * z,
    dt = 2e-8 s.  Force_State.cu:29-54,203-306 arithmetic.
    """
    dx = 0.5e-3
    p, nxyz, _ = params(3, _lib.PD_MODEL_LPS, 2.0e-8, dx, length=nx * dx, height=ny * dx, width=nz * dx,
                        emod=70e9, pratio=0.33, dens=2700.0, s0=0.002)
    p.precision = precision
    x, y, z = gen_block_jitter(nxyz[0], nxyz[1], nxyz[2], dx, 0.01, seed)
    v0 = np.zeros((3, len(x)))
    v0[2] = 1.0e3 * z

    def setup(prob):
        prob.init_state()
        prob.lps_weights()

    return Scenario("c2_lps_block", p, x, y, z, len(x), steps, nxyz, setup, v0=v0)


def pmb_block(nx=200, ny=200, nz=100, steps=200, precision=_lib.PD_PREC_MIXED, seed=0x5EED0002):
    """PMB force on the same jittered 4M-particle block (E = 70 GPa, nu = 0.25)."""
    dx = 0.5e-3
    p, nxyz, _ = params(3, _lib.PD_MODEL_PMB, 2.0e-8, dx, thick=dx, length=nx * dx, height=ny * dx,
                        width=nz * dx, emod=70e9, pratio=0.25, dens=2700.0, s0=0.002)
    p.precision = precision
    x, y, z = gen_block_jitter(nxyz[0], nxyz[1], nxyz[2], dx, 0.01, seed)
    v0 = np.zeros((3, len(x)))
    v0[2] = 1.0e3 * z

    def setup(prob):
        prob.init_state()

    return Scenario("pmb_block", p, x, y, z, len(x), steps, nxyz, setup, v0=v0)


def plate_tensile(dx=0.01, nt=1000):
    """2D PMB plate under edge body forces solved by ADR (Solve_Plate_Tensile.cu:30-137).

    1.0 x 0.5 m plate, t = 0.01, E = 200 GPa, nu = 1/3, rho = 7850, s0 = 1.0,
    fictitious dt = 1, body force +-0.001 E/dx on the edge columns.
    """
    p, nxyz, _ = params(2, _lib.PD_MODEL_PMB, 1.0, dx, thick=0.01, length=1.0, height=0.5, emod=200e9,
                        pratio=1.0 / 3.0, dens=7850.0, s0=1.0)
    nx, ny, _ = nxyz
    x, y = gen_plate(nx, ny, dx)

    def setup(prob):
        prob.surface_correct()
        prob.init_state(static=True)
        prob.set_body_force_tensile(nx, ny)

    return Scenario("plate_tensile", p, x, y, None, nx * ny, nt, nxyz, setup)


def params_dict(p):
    """Plain dict of a PdParams (used to drive the FP64 oracle with the same scalars)."""
    return {name: getattr(p, name) for name, _ in p._fields_}
