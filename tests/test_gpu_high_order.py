"""Orders above the reference's operator files (p = 16..19): operators from
the closed-form generator, stage kernels on the direction-by-direction MFMA
path, against the oracle run on the same generated operators (written in the
reference's file format).

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


def test_high_order_piecewise_path_is_refused():
    p = 17
    with dgk.Solver.structured("HIP", 1, 1, 2, p) as s:
        q, u, v, w = inputs.random_fields(1, 1, 2, p, seed=5)
        s.set_fields(q, u, v, w)
        flux = s.cal_elembnd_flux()
        with pytest.raises(dgk.DgkError) as ei:
            s.cal_dqdt(flux)
        assert ei.value.code == dgk.ERR_UNSUPPORTED
