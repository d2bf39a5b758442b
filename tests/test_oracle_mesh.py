"""Pins the oracle's restatement of mesh_setup (mod_mesh.f90:74-193) against
geometric invariants, and checks the product's host mesh builder against it
(CPU).

Invariants (independent of the map construction):
  * every face node and the node VMapP (through the halo) points to sit at the
    same physical position modulo the unit period;
  * VMapM nodes lie on the element's face f, whose outward normal is
    normal_fn; Escale = 2/h and Fscale = 2/h_normal;
  * boundary VMapP entries are Np*Ne+1 .. Np*Ne+NhaloNode, handed out in the
    serialized (k, j, i, f, fp) order, and halo_src_map addresses owned nodes;
  * update_halo copies exactly halo_src_map's sources.
"""
import numpy as np
import pytest

import dgk
import oracle

MESHES = [(1, 1, 1, 1), (2, 3, 2, 3), (3, 2, 4, 2), (4, 4, 4, 1), (2, 2, 2, 7), (1, 2, 3, 5)]


def positions(m):
    """pos(NeA*Np flat, 3) with the halo tail taken from its source nodes."""
    pos = m.pos_en.transpose(1, 2, 0).reshape(-1, 3)  # (Ne*Np, 3)
    halo = pos[m.halo_src_map - 1]
    return np.concatenate([pos, halo])


@pytest.mark.parametrize("mesh", MESHES)
def test_face_pairs_meet(mesh):
    m = oracle.Mesh(*mesh)
    pos = positions(m)
    pM = pos[m.VMapM.reshape(-1) - 1]
    pP = pos[m.VMapP.reshape(-1) - 1]
    d = pM - pP
    d -= np.rint(d)  # unit period
    assert np.abs(d).max() <= 1e-14


@pytest.mark.parametrize("mesh", MESHES)
def test_face_geometry(mesh):
    nx, ny, nz, p = mesh
    m = oracle.Mesh(*mesh)
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    axis = [1, 0, 1, 0, 2, 2]
    sign = [-1, 1, 1, -1, -1, 1]
    pos = m.pos_en.transpose(1, 2, 0)  # (Ne, Np, 3)
    for e in range(m.Ne):
        lo, hi = pos[e].min(axis=0), pos[e].max(axis=0)
        for f in range(6):
            rows = slice(f * m.Nfp, (f + 1) * m.Nfp)
            nodes = m.VMapM[e, rows] - 1 - e * m.Np
            assert nodes.min() >= 0 and nodes.max() < m.Np
            coord = pos[e, nodes, axis[f]]
            want = hi[axis[f]] if sign[f] > 0 else lo[axis[f]]
            assert np.abs(coord - want).max() <= 1e-15
            nrm = np.zeros(3)
            nrm[axis[f]] = sign[f]
            assert np.array_equal(m.normal_fn[:, e, rows].T, np.tile(nrm, (m.Nfp, 1)))
            assert np.all(m.Fscale[e, rows] == 2.0 / h[axis[f]])
        for d in range(3):
            assert np.all(m.Escale[d, e] == 2.0 / h[d])


@pytest.mark.parametrize("mesh", MESHES)
def test_halo_order_and_update(mesh):
    m = oracle.Mesh(*mesh)
    owned = m.Np * m.Ne
    vp = m.VMapP.reshape(-1)  # (e, f, fp) order == (k, j, i, f, fp) order
    bnd = vp[vp > owned]
    assert np.array_equal(bnd, owned + 1 + np.arange(m.NhaloNode))
    assert m.halo_src_map.min() >= 1 and m.halo_src_map.max() <= owned
    q = np.random.default_rng(0).standard_normal((m.Ne, m.Np))
    f = m.field(q).reshape(-1)
    assert np.array_equal(f[owned:owned + m.NhaloNode], q.reshape(-1)[m.halo_src_map - 1])
    assert m.NhaloNode == m.Nfp * 2 * (mesh[1] * mesh[2] + mesh[0] * mesh[2] + mesh[0] * mesh[1])


@pytest.mark.parametrize("mesh", MESHES)
def test_product_mesh_builder_matches_oracle(mesh):
    m = oracle.Mesh(*mesh)
    pm = dgk.mesh_build(*mesh, m.x1D)
    for k in ("VMapM", "VMapP", "halo_src_map", "normal_fn", "Escale", "Fscale", "pos_en"):
        assert np.array_equal(getattr(m, k), pm[k]), k
    assert (pm["Ne"], pm["NeA"], pm["NhaloNode"]) == (m.Ne, m.NeA, m.NhaloNode)
