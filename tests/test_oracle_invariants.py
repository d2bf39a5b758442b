"""Pins the oracle's tendency and time loop by what the mathematics fixes (CPU).

* A dense brute-force tendency built independently of the oracle's maps and
  loops: neighbours found by coordinates, derivatives as Kronecker products of
  the LGL matrix, the Lift from the closed-form 1D lift, and the numerical flux
  written as Fscale * (F* - F_M) with the Lax-Friedrichs flux
  F* = {F.n} - alpha/2 [q] (Hesthaven & Warburton 2008, sec. 2.2).
* Discrete conservation (w^T D = e_N - e_1 and a single-valued F*), free-stream
  preservation, energy decay for constant velocity (upwind), and the SSP-RK3
  amplification factor 1 + z + z^2/2 + z^3/6 (Gottlieb & Shu 1998).
"""
import numpy as np
import pytest

import oracle
from dgk import inputs
from test_oracle_operators import diff_matrix, ell, lgl


def brute_tendency(nx, ny, nz, p, q, u, v, w):
    """dqdt (Ne, Np) for the periodic unit box, from first principles."""
    n = p + 1
    Np = n ** 3
    x, wq, Pp = lgl(p)
    D = diff_matrix(p, x, Pp)
    I = np.eye(n)
    Dx = np.kron(I, np.kron(I, D))  # node index i + j n + k n^2: i fastest
    Dy = np.kron(I, np.kron(D, I))
    Dz = np.kron(D, np.kron(I, I))
    lm, lp = ell(p, x, -1), ell(p, x, +1)
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    ne = np.array([nx, ny, nz])
    idx = np.arange(Np)
    a = np.stack([idx % n, (idx // n) % n, idx // (n * n)], axis=1)  # (Np, 3) node coords
    out = np.zeros((nx * ny * nz, Np))
    vel = np.stack([u, v, w], axis=-1)  # (Ne, Np, 3)
    for e in range(nx * ny * nz):
        ec = np.array([e % nx, (e // nx) % ny, e // (nx * ny)])
        F = q[e][:, None] * vel[e]  # (Np, 3)
        vol = (2.0 / h[0]) * (Dx @ F[:, 0]) + (2.0 / h[1]) * (Dy @ F[:, 1]) + (2.0 / h[2]) * (Dz @ F[:, 2])
        surf = np.zeros(Np)
        for axis in range(3):
            for side in (-1, 1):
                on_face = a[:, axis] == (0 if side < 0 else n - 1)
                nb = ec.copy()
                nb[axis] = (nb[axis] + side) % ne[axis]
                enb = nb[0] + nb[1] * nx + nb[2] * nx * ny
                # the neighbour's node at the same physical point: same
                # tangential indices, opposite end along the axis
                anb = a[on_face].copy()
                anb[:, axis] = n - 1 if side < 0 else 0
                inb = anb[:, 0] + anb[:, 1] * n + anb[:, 2] * n * n
                qM, qP = q[e][on_face], q[enb][inb]
                VM = side * vel[e][on_face, axis]
                VP = side * vel[enb][inb, axis]
                alpha = 0.5 * np.abs(VM + VP)
                Fstar = 0.5 * (qM * VM + qP * VP) - 0.5 * alpha * (qP - qM)
                flux = (2.0 / h[axis]) * (Fstar - qM * VM)
                lift1d = lm if side < 0 else lp
                # face node with the same tangential indices as node `idx`
                key = {tuple(np.delete(a[m], axis)): k for k, m in enumerate(np.nonzero(on_face)[0])}
                for node in range(Np):
                    t = tuple(np.delete(a[node], axis))
                    surf[node] += lift1d[a[node, axis]] * flux[key[t]]
        out[e] = -(vol + surf)
    return out


@pytest.mark.parametrize("mesh", [(1, 1, 1, 1), (2, 1, 1, 2), (2, 2, 1, 3), (1, 2, 2, 4), (2, 2, 2, 2)])
def test_tendency_equals_dense_brute_force(mesh):
    nx, ny, nz, p = mesh
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=31 + p)
    m = oracle.Mesh(nx, ny, nz, p)
    got = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    ref = brute_tendency(nx, ny, nz, p, q, u, v, w)
    # the brute force uses closed-form operators (data files agree to ~2e-12)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def quadrature_weights(p, nx, ny, nz):
    _, wq, _ = lgl(p)
    w3 = np.einsum("k,j,i->kji", wq, wq, wq).reshape(-1)
    return w3 * (1.0 / (8 * nx * ny * nz))


def test_discrete_conservation_any_velocity():
    nx, ny, nz, p = 3, 2, 2, 4
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=41, vel_scale=0.7)
    m = oracle.Mesh(nx, ny, nz, p)
    W = quadrature_weights(p, nx, ny, nz)
    mass0 = (q * W).sum()
    qf, _, _ = m.run(q, u, v, w, 2e-3, 20)
    assert abs((qf * W).sum() - mass0) <= 1e-13 * np.abs(q * W).sum()


def test_free_stream_is_preserved():
    nx, ny, nz, p = 2, 3, 2, 5
    m = oracle.Mesh(nx, ny, nz, p)
    one = np.full((m.Ne, m.Np), 1.7)
    vel = [np.full((m.Ne, m.Np), c) for c in (0.3, -0.8, 0.55)]
    dq = m.cal_tend(m.field(one), *(m.field(x) for x in vel))
    assert np.abs(dq).max() <= 1e-11


def test_energy_decays_for_constant_velocity():
    wl = inputs.c1_sine_noise(ne=4, p=4, nstep=30)
    m = oracle.Mesh(4, 4, 4, 4)
    W = quadrature_weights(4, 4, 4, 4)
    e_prev = (wl.q ** 2 * W).sum()
    q = wl.q
    for _ in range(5):
        q, _, _ = m.run(q, wl.u, wl.v, wl.w, wl.dt, 6)
        e = (q ** 2 * W).sum()
        assert e <= e_prev * (1 + 1e-14)
        e_prev = e


def test_ssp_rk3_amplification_factor():
    """Stage updates (main.f90:69-73) with dqdt = lam*q reproduce the
    third-order SSP-RK3 stability polynomial."""
    n = 257
    rng = np.random.default_rng(5)
    q0 = rng.standard_normal(n)
    for z in (-0.3, -1.1, 0.45):
        lam, dt = z / 0.01, 0.01
        q = q0.copy()
        for stage in range(3):
            dq = lam * q
            oracle.lib().orc_rk_update(oracle._p(q), oracle._p(q0), oracle._p(dq), dt, stage, 1, n)
        amp = 1 + z + z * z / 2 + z ** 3 / 6
        assert np.abs(q - amp * q0).max() <= 1e-14 * np.abs(q0).max()


def _translation_error(p, ne=4, dt=1e-3, nstep=100):
    m = oracle.Mesh(ne, ne, ne, p)
    X, Y, Z = (m.pos_en[d] for d in range(3))
    prof = lambda x: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * Y) + 0.3 * np.sin(2 * np.pi * Z)
    u = np.ones_like(X)
    zero = np.zeros_like(X)
    q, _, _ = m.run(prof(X), u, zero, zero, dt, nstep)
    return np.abs(q - prof(X - dt * nstep)).max()


def test_translation_of_a_smooth_profile():
    """Constant velocity (1,0,0): q(x,t) = q0(x - t) exactly; the error on
    4^3 elements falls spectrally with p."""
    e4, e6 = _translation_error(4), _translation_error(6)
    assert e6 <= 1e-5
    assert e6 <= e4 / 10
