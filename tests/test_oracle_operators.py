"""Pins the operator data read by the oracle and the device path against
closed forms (CPU).

operator_data/p{p}.dat holds x1D, D1D and Lift_mat (element_init,
mod_mesh.f90:253-261).  Independently of the files:
  * x1D are the Legendre-Gauss-Lobatto nodes: +-1 and the roots of P_p';
  * D1D is the LGL differentiation matrix D_ij = P_p(x_i) / (P_p(x_j)(x_i - x_j)),
    D_00 = -p(p+1)/4, D_pp = p(p+1)/4, other diagonal entries 0
    (Kopriva 2009, ch. 3; Hesthaven & Warburton 2008, sec. 3.2);
  * Lift_mat(:,:,:,f) = I (x) l-/+ (x) I along the face normal, with
    l-/+(x) = sum_n (2n+1)/2 P_n(x) P_n(-/+1), the exact-mass-matrix lift;
  * identities: D 1 = 0, D x = 1, W D + D^T W = diag(-1, 0, ..., 0, 1),
    w^T l-/+ = 1 with w the LGL weights 2/(p(p+1) P_p(x_i)^2).
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

import oracle

ORDERS = range(1, 16)


def lgl(p):
    c = np.zeros(p + 1)
    c[-1] = 1.0
    inner = L.legroots(L.legder(c)) if p > 1 else np.array([])
    x = np.concatenate(([-1.0], np.sort(inner), [1.0]))
    Pp = L.legval(x, c)
    w = 2.0 / (p * (p + 1) * Pp ** 2)
    return x, w, Pp


def diff_matrix(p, x, Pp):
    n = p + 1
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = Pp[i] / (Pp[j] * (x[i] - x[j]))
    D[0, 0] = -p * (p + 1) / 4.0
    D[-1, -1] = p * (p + 1) / 4.0
    return D


def ell(p, x, side):
    s = -1.0 if side < 0 else 1.0
    out = np.zeros_like(x)
    for n in range(p + 1):
        c = np.zeros(n + 1)
        c[-1] = 1.0
        out += (2 * n + 1) / 2.0 * L.legval(x, c) * L.legval(s, c)
    return out


@pytest.mark.parametrize("p", ORDERS)
def test_nodes_and_differentiation_matrix(p):
    x1D, Dm, _ = oracle.load_operators(p)
    D = Dm.T  # D[i, l] = D1D(i, l)
    x, w, Pp = lgl(p)
    assert np.abs(x1D - x).max() <= 2e-15
    Dref = diff_matrix(p, x, Pp)
    assert np.abs(D - Dref).max() <= 5e-12
    n = p + 1
    scale = np.abs(D).max()
    assert np.abs(D @ np.ones(n)).max() <= 1e-13 * scale
    assert np.abs(D @ x1D - 1.0).max() <= 1e-13 * scale
    W = np.diag(w)
    Q = W @ D + D.T @ W
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.abs(Q - B).max() <= 1e-13 * scale
    assert np.abs(w @ D - (np.eye(n)[-1] - np.eye(n)[0])).max() <= 1e-13 * scale


@pytest.mark.parametrize("p", ORDERS)
def test_lift_is_tensor_product_of_exact_1d_lift(p):
    _, _, Lm = oracle.load_operators(p)  # Lm[f, k, j, i] = Lift_mat(i, j, k, f)
    x, w, _ = lgl(p)
    lm, lp = ell(p, x, -1), ell(p, x, +1)
    n = p + 1
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # faces -y, +x, +y, -x, -z, +z: 1D lift along y, x, y, x, z, z
    along = [J, I, J, I, K, K]
    side = [lm, lp, lp, lm, lm, lp]
    scale = np.abs(Lm).max()
    for f in range(6):
        ref = side[f][along[f]]           # indexed [i, j, k]
        got = Lm[f].transpose(2, 1, 0)    # -> [i, j, k]
        assert np.abs(got - ref).max() <= 5e-12 * scale, f
    assert abs(w @ lm - 1.0) <= 1e-13 * np.abs(lm).max()
    assert abs(w @ lp - 1.0) <= 1e-13 * np.abs(lp).max()


def test_operator_files_match_product_loader():
    import dgk

    for p in (1, 7, 15):
        a = oracle.load_operators(p)
        b = dgk.load_operators(p)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
