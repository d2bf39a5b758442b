"""Deterministic synthetic inputs for the advection workloads.

Random numbers come from splitmix64 (S. Vigna, public domain) applied to a
counter: u01(seed, c) = (splitmix64((seed << 40) ^ c) >> 11) * 2**-53, so
every component (this module, the C oracle, the serving generator) produces
bit-identical inputs from a seed.  Per-node draws use the global DOF index
c = e*Np + n; scalar parameters use counters from PARAM_BASE upwards.

Node coordinates follow set_element_position (mod_mesh.f90:332-353) on the
periodic box [0,Lx]x[0,Ly]x[0,Lz] (mesh_setup, mod_mesh.f90:112-143).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import load_operators

PARAM_BASE = 1 << 39
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def u01(seed: int, counter) -> np.ndarray:
    c = np.asarray(counter, dtype=np.uint64)
    x = (np.uint64(seed) << np.uint64(40)) ^ c
    return (splitmix64(x) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def node_coords(NeX, NeY, NeZ, x1D, dom=(1.0, 1.0, 1.0)):
    """(X, Y, Z), each of numpy shape (Ne, Np) = the reference's pos_en(Np,Ne,d)."""
    Nq = len(x1D)
    h = [dom[0] / NeX, dom[1] / NeY, dom[2] / NeZ]
    ie = np.arange(NeX * NeY * NeZ)
    ei = [ie % NeX, (ie // NeX) % NeY, ie // (NeX * NeY)]
    n = np.arange(Nq ** 3)
    ni = [n % Nq, (n // Nq) % Nq, n // (Nq * Nq)]
    out = []
    for d in range(3):
        c = ((ei[d] + 1).astype(np.float64) - 0.5) * h[d]
        out.append(c[:, None] + (0.5 * h[d]) * np.asarray(x1D)[ni[d]][None, :])
    return tuple(out)


def cfl_dt(x1D, Ne_axes, vel, cfl=0.1, dom=(1.0, 1.0, 1.0)) -> float:
    """dt = cfl * h_min / (|u|+|v|+|w|), h_min = min_d (x1D[1]-x1D[0]) * (L_d/Ne_d) / 2."""
    gap = float(x1D[1] - x1D[0])
    hmin = min(gap * (dom[d] / Ne_axes[d]) / 2.0 for d in range(3))
    speed = sum(abs(float(c)) for c in vel)
    return cfl * hmin / speed


def gaussian_blob(X, Y, Z, centre, sigma):
    """exp(-r^2 / (2 sigma^2)), r the periodic minimum-image distance in [0,1)^3."""
    r2 = np.zeros_like(X)
    for Xd, cd in zip((X, Y, Z), centre):
        d = Xd - cd
        d = d - np.rint(d)
        r2 += d * d
    return np.exp(-r2 / (2.0 * sigma * sigma))


@dataclass
class Workload:
    name: str
    NeX: int
    NeY: int
    NeZ: int
    p: int
    vel: tuple
    nstep: int
    dt: float
    q: np.ndarray = field(repr=False)       # (Ne, Np)
    u: np.ndarray = field(repr=False)
    v: np.ndarray = field(repr=False)
    w: np.ndarray = field(repr=False)

    @property
    def Ne(self) -> int:
        return self.NeX * self.NeY * self.NeZ

    @property
    def Np(self) -> int:
        return (self.p + 1) ** 3

    @property
    def dof(self) -> int:
        return self.Ne * self.Np


def _const_vel(shape, vel):
    return tuple(np.full(shape, float(c)) for c in vel)


def c1_sine_noise(ne=8, p=3, seed=1, nstep=100) -> Workload:
    """8^3 elements at p=3, velocity (1.0, 0.5, 0.25),
    q = sin(2 pi x) sin(2 pi y) sin(2 pi z) + 1e-3 U(-1,1) (seed 1)."""
    x1D, _, _ = load_operators(p)
    X, Y, Z = node_coords(ne, ne, ne, x1D)
    pi = 4.0 * np.arctan(1.0)
    Ne, Np = X.shape
    noise = 2.0 * u01(seed, np.arange(Ne * Np, dtype=np.uint64)).reshape(Ne, Np) - 1.0
    q = np.sin(2.0 * pi * X) * np.sin(2.0 * pi * Y) * np.sin(2.0 * pi * Z) + 1e-3 * noise
    vel = (1.0, 0.5, 0.25)
    u, v, w = _const_vel(q.shape, vel)
    return Workload("c1_8cube_p3", ne, ne, ne, p, vel, nstep, cfl_dt(x1D, (ne,) * 3, vel),
                    q, u, v, w)


def c2_gaussian(ne=32, p=7, seed=2, nstep=200, sigma=0.1) -> Workload:
    """32^3 elements at p=7, velocity components U[-1,1] and a Gaussian blob
    (sigma 0.1) centred at a U[0,1)^3 point, all from one seed."""
    x1D, _, _ = load_operators(p)
    par = u01(seed, PARAM_BASE + np.arange(6, dtype=np.uint64))
    vel = tuple(float(2.0 * par[c] - 1.0) for c in range(3))
    centre = tuple(float(par[3 + c]) for c in range(3))
    X, Y, Z = node_coords(ne, ne, ne, x1D)
    q = gaussian_blob(X, Y, Z, centre, sigma)
    del X, Y, Z
    u, v, w = _const_vel(q.shape, vel)
    return Workload(f"c2_{ne}cube_p{p}_gauss", ne, ne, ne, p, vel, nstep,
                    cfl_dt(x1D, (ne,) * 3, vel), q, u, v, w)


def c3_sine(ne=16, p=15, nstep=100, wavenumber=3) -> Workload:
    """16^3 elements at p=15, velocity (1,1,1), q = prod sin(2 pi k x_d), k = 3."""
    x1D, _, _ = load_operators(p)
    X, Y, Z = node_coords(ne, ne, ne, x1D)
    pi = 4.0 * np.arctan(1.0)
    kk = 2.0 * pi * wavenumber
    q = np.sin(kk * X) * np.sin(kk * Y) * np.sin(kk * Z)
    vel = (1.0, 1.0, 1.0)
    u, v, w = _const_vel(q.shape, vel)
    return Workload(f"c3_{ne}cube_p{p}_sine", ne, ne, ne, p, vel, nstep,
                    cfl_dt(x1D, (ne,) * 3, vel), q, u, v, w)


def random_fields(NeX, NeY, NeZ, p, seed, vel_scale=1.0):
    """Random q and per-node random velocity fields (U(-1,1) each), for
    tendency parity at arbitrary operator inputs."""
    Ne, Np = NeX * NeY * NeZ, (p + 1) ** 3
    n = Ne * Np
    c = np.arange(n, dtype=np.uint64)
    q = 2.0 * u01(seed, c) - 1.0
    u = vel_scale * (2.0 * u01(seed, c + np.uint64(n)) - 1.0)
    v = vel_scale * (2.0 * u01(seed, c + np.uint64(2 * n)) - 1.0)
    w = vel_scale * (2.0 * u01(seed, c + np.uint64(3 * n)) - 1.0)
    return (q.reshape(Ne, Np), u.reshape(Ne, Np), v.reshape(Ne, Np), w.reshape(Ne, Np))


def ensemble(n=512, seed=4, ne_choices=(4, 6, 8, 12, 16), p_choices=(3, 5, 7, 11),
             sigma_range=(0.05, 0.2), nstep=50, cfl=0.1):
    """Batched-ensemble instances: instance k draws, from counters of seed
    `seed` starting at PARAM_BASE + 16k, Ne per axis and p uniformly from the
    given choices, a velocity with components U[-1,1], a Gaussian-blob centre
    U[0,1)^3 and sigma U[sigma_range]; dt from CFL `cfl`."""
    out = []
    for k in range(n):
        r = u01(seed, PARAM_BASE + 16 * k + np.arange(9, dtype=np.uint64))
        ne = int(ne_choices[min(int(r[0] * len(ne_choices)), len(ne_choices) - 1)])
        p = int(p_choices[min(int(r[1] * len(p_choices)), len(p_choices) - 1)])
        vel = tuple(float(2.0 * r[2 + c] - 1.0) for c in range(3))
        centre = tuple(float(r[5 + c]) for c in range(3))
        sigma = float(sigma_range[0] + (sigma_range[1] - sigma_range[0]) * r[8])
        x1D = lgl_nodes(p)
        out.append(dict(ne=(ne, ne, ne), p=p, vel=vel, centre=centre, sigma=sigma,
                        nstep=nstep, dt=cfl_dt(x1D, (ne, ne, ne), vel, cfl=cfl)))
    return out


_LGL = {}


def lgl_nodes(p):
    """x1D of operator_data/p{p}.dat (cached)."""
    if p not in _LGL:
        _LGL[p] = load_operators(p)[0]
    return _LGL[p]


def serve_requests(n=10000, seed=5, rate=4.0, ne_range=(4, 32), p_range=(2, 15),
                   nstep_range=(10, 500), sigma_range=(0.05, 0.2), mp_fraction=0.5):
    """Serving stream: request k draws, from counters of seed `seed` starting at
    PARAM_BASE + 16k, Ne per axis independently and uniformly from ne_range
    (inclusive), p from p_range, a velocity with components U[-1,1], a
    Gaussian-blob centre U[0,1)^3, sigma U[sigma_range], nstep from
    nstep_range and the tolerance class (1e-5 with probability mp_fraction,
    else FP64); its arrival follows request k-1 after an exponential gap of
    mean 1/rate seconds (a Poisson stream of rate `rate`).  dt is left to the
    server (CFL 0.1)."""
    out = []
    t = 0.0

    def pick(u, lo, hi):
        return int(lo + min(int(u * (hi - lo + 1)), hi - lo))

    for k in range(n):
        r = u01(seed, PARAM_BASE + 16 * k + np.arange(14, dtype=np.uint64))
        t += -np.log1p(-float(r[13])) / rate
        out.append(dict(
            ne=tuple(pick(float(r[c]), *ne_range) for c in range(3)),
            p=pick(float(r[3]), *p_range),
            vel=tuple(float(2.0 * r[4 + c] - 1.0) for c in range(3)),
            centre=tuple(float(r[7 + c]) for c in range(3)),
            sigma=float(sigma_range[0] + (sigma_range[1] - sigma_range[0]) * r[10]),
            nstep=pick(float(r[11]), *nstep_range),
            tol_class=1 if float(r[12]) < mp_fraction else 0,
            arrival=float(t)))
    return out
