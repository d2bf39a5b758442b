"""Orders above the reference's operator files (p = 16..19): operators from
the closed-form generator, stage kernels on the MFMA path (one element per
workgroup, all at once or direction by direction), against the oracle run on
the same generated operators (written in the reference's file format).

FP64: tendency max-rel <= 1e-13 and 3-step loop max-rel <= 1e-12, through
both creation routes (reference arrays and structured box).  Mixed
precision: 1e-5 relative L2 after the loop.
"""
import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

pytestmark = pytest.mark.gpu

TEND_TOL = 1e-13
LOOP_TOL = 1e-12
MP_TOL = 1e-5


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("p", [16, 17, 18, 19])
def test_high_order_against_oracle(tmp_path, p):
    dgk.write_operators(tmp_path, p)
    ne = (2, 1, 2)
    m = oracle.Mesh(*ne, p, operator_dir=tmp_path)
    q, u, v, w = inputs.random_fields(*ne, p, seed=1100 + p, vel_scale=0.7)
    x1D = dgk.make_operators(p)[0]
    dt = 0.2 * inputs.cfl_dt(x1D, ne, (0.7, 0.7, 0.7))
    tend_ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    loop_ref, _, _ = m.run(q, u, v, w, dt, 3)
    with dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, m.D, m.Lift, m.VMapM,
                           m.VMapP, m.halo_src_map, m.normal_fn, m.Escale, m.Fscale) as s:
        s.set_fields(q, u, v, w)
        assert maxrel(s.cal_tend(), tend_ref) <= TEND_TOL
        s.step(dt, 3)
        assert maxrel(s.get_field(dgk.FIELD_Q), loop_ref) <= LOOP_TOL
    with dgk.Solver.structured("HIP", *ne, p) as s:  # operators generated in the library
        s.set_fields(q, u, v, w)
        s.step(dt, 3)
        assert maxrel(s.get_field(dgk.FIELD_Q), loop_ref) <= LOOP_TOL
    with dgk.Solver.structured("HIP_MP", *ne, p) as s:
        s.set_fields(q, u, v, w)
        s.step(dt, 3)
        got = s.get_field(dgk.FIELD_Q)
    assert np.linalg.norm(got - loop_ref) <= MP_TOL * np.linalg.norm(loop_ref)


@pytest.mark.parametrize("p", [16, 17])
def test_high_order_fp64_shapes_against_oracle(tmp_path, p):
    """Both FP64 shapes where the all-at-once one fits (p = 16, 17) hold the
    tendency and loop bars and give the same bits."""
    dgk.write_operators(tmp_path, p)
    ne = (2, 1, 3)
    m = oracle.Mesh(*ne, p, operator_dir=tmp_path)
    q, u, v, w = inputs.random_fields(*ne, p, seed=1200 + p, vel_scale=0.8)
    x1D = dgk.make_operators(p)[0]
    dt = 0.2 * inputs.cfl_dt(x1D, ne, (0.8, 0.8, 0.8))
    tend_ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    loop_ref, _, _ = m.run(q, u, v, w, dt, 3)
    assert dgk.tile_shapes("HIP", p) == [1, 1]
    loops = []
    try:
        for cand in range(2):
            dgk.tile_select("HIP", p, cand)
            with dgk.Solver.structured("HIP", *ne, p) as s:
                s.set_fields(q, u, v, w)
                assert maxrel(s.cal_tend(), tend_ref) <= TEND_TOL, cand
                s.step(dt, 3)
                got = s.get_field(dgk.FIELD_Q)
            assert maxrel(got, loop_ref) <= LOOP_TOL, cand
            loops.append(got)
    finally:
        dgk.tile_select("HIP", p, -1)
    assert np.array_equal(loops[1], loops[0])


def test_high_order_piecewise_path_is_refused():
    p = 17
    with dgk.Solver.structured("HIP", 1, 1, 2, p) as s:
        q, u, v, w = inputs.random_fields(1, 1, 2, p, seed=5)
        s.set_fields(q, u, v, w)
        flux = s.cal_elembnd_flux()
        with pytest.raises(dgk.DgkError) as ei:
            s.cal_dqdt(flux)
        assert ei.value.code == dgk.ERR_UNSUPPORTED
