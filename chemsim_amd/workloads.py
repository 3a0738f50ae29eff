"""Initial states and geometries of the benchmark workloads.

Orientation (src/lbm.rs:716-729 with the stencils of :233-269): population i
moves by (drow, dcol) = (-cx_i, +cy_i) per step.  A position is therefore
(X, Y) = (-r, c) in lattice-velocity coordinates: the velocity component ux
points toward decreasing row, uy toward increasing column.  Row 0 is the
image top (src/render.rs:15-19).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

FLUID, WALL, LID, INLET, OUTLET = range(5)


def viscosity(tau: float) -> float:
    """Kinematic shear viscosity of BGK, nu = cs^2 (tau - dt/2) (src/lbm.rs:366-369), dx = dt = 1."""
    return (tau - 0.5) / 3.0


def tau_for_viscosity(nu: float) -> float:
    return 3.0 * nu + 0.5


# ------------------------------------------------------------------ cavity --

def cavity_flags(w: int, h: int) -> np.ndarray:
    """Lid-driven cavity: row 0 is the moving lid, the other three edges are walls."""
    f = np.zeros((h, w), np.uint8)
    f[-1, :] = WALL
    f[:, 0] = WALL
    f[:, -1] = WALL
    f[0, 1:-1] = LID
    f[0, 0] = WALL
    f[0, -1] = WALL
    return f


def lid_velocity(u: float) -> tuple[float, float]:
    """Lid moving toward +column: lattice velocity (ux, uy) = (0, u)."""
    return (0.0, float(u))


def rest_state(w: int, h: int, rho: float = 1.0):
    """rho = 1, u = 0 everywhere: f = f_eq(1, 0) = w_i (src/main.rs:257-265 with zero velocity)."""
    return (np.full((h, w), rho, np.float32), np.zeros((h, w), np.float32), np.zeros((h, w), np.float32))


# ------------------------------------------------------------ Taylor-Green --

def taylor_green(n: int, u0: float, dtype=np.float64, m: int | None = None):
    """Periodic Taylor-Green vortex on an n x n (or n x m) lattice, wavenumber 2 pi / n.

    ux = -u0 cos(kX) sin(kY), uy = u0 sin(kX) cos(kY),
    rho = 1 - (3 u0^2 / 4)(cos 2kX + cos 2kY), with (X, Y) = (-r, c).
    Decays as exp(-2 nu k^2 t).
    """
    h = n
    w = n if m is None else m
    k = 2.0 * math.pi / n
    r = np.arange(h, dtype=np.float64)[:, None]
    c = np.arange(w, dtype=np.float64)[None, :]
    X, Y = -r, c
    ux = -u0 * np.cos(k * X) * np.sin(k * Y)
    uy = u0 * np.sin(k * X) * np.cos(k * Y)
    rho = 1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    ux = np.broadcast_to(ux, (h, w))
    uy = np.broadcast_to(uy, (h, w))
    rho = np.broadcast_to(rho, (h, w))
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (rho, ux, uy))


def taylor_green_decay(n: int, tau: float, steps: int) -> float:
    k = 2.0 * math.pi / n
    return math.exp(-2.0 * viscosity(tau) * k * k * steps)


# ----------------------------------------------------------------- channel --

@dataclass
class Channel:
    """Channel with a Zou-He velocity inlet (column 0), density outlet (last column),
    walls on the first and last rows and circular obstacles."""
    w: int
    h: int
    circles: list = field(default_factory=list)   # (row, col, radius)
    re: float = 50.0
    u_in: float = 0.05

    @property
    def width(self) -> float:
        """Fluid width between the half-way walls."""
        return float(self.h - 2)

    @property
    def nu(self) -> float:
        """Viscosity from Re = u_in * width / nu (Reynolds number on the channel width)."""
        return self.u_in * self.width / self.re

    @property
    def tau(self) -> float:
        return tau_for_viscosity(self.nu)

    def mask(self) -> np.ndarray:
        """Obstacle mask (1 = solid), h x w."""
        m = np.zeros((self.h, self.w), np.uint8)
        rr = np.arange(self.h)[:, None]
        cc = np.arange(self.w)[None, :]
        for (r0, c0, rad) in self.circles:
            m[(rr - r0) ** 2 + (cc - c0) ** 2 <= rad * rad] = 1
        return m

    def flags(self) -> np.ndarray:
        f = np.zeros((self.h, self.w), np.uint8)
        f[self.mask() == 1] = WALL
        f[1:-1, 0] = INLET
        f[1:-1, -1] = OUTLET
        f[0, :] = WALL
        f[-1, :] = WALL
        return f

    def inlet(self) -> tuple[float, float]:
        """Inlet velocity along +column: (ux, uy) = (0, u_in)."""
        return (0.0, float(self.u_in))


def random_channel(rng: np.random.Generator, w: int = 512, h: int = 128, u_in: float = 0.05,
                   re_range=(20.0, 100.0)) -> Channel:
    """1-4 circular obstacles of radius 6-20 cells at seeded positions, Re uniform in re_range."""
    n = int(rng.integers(1, 5))
    circles = []
    tries = 0
    while len(circles) < n and tries < 1000:
        tries += 1
        rad = int(rng.integers(6, 21))
        r0 = int(rng.integers(1 + rad + 2, h - 1 - rad - 2))
        c0 = int(rng.integers(rad + 24, w - rad - 96))
        if all((r0 - a) ** 2 + (c0 - b) ** 2 > (rad + q + 4) ** 2 for (a, b, q) in circles):
            circles.append((r0, c0, rad))
    re = float(rng.uniform(*re_range))
    return Channel(w=w, h=h, circles=circles, re=re, u_in=u_in)


# ------------------------------------------------------------ batch config --

@dataclass
class CavitySpec:
    side: int
    tau: float
    u_lid: float


def batch_cavities(seed: int = 0, count: int = 512, lo: int = 64, hi: int = 512) -> list[CavitySpec]:
    """count independent lid-driven cavities: side ~ U{lo..hi}, tau ~ U[0.55, 1.0], u_lid ~ U[0.02, 0.1]."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=count)
    taus = rng.uniform(0.55, 1.0, size=count)
    lids = rng.uniform(0.02, 0.1, size=count)
    return [CavitySpec(int(s), float(t), float(u)) for s, t, u in zip(sides, taus, lids)]
