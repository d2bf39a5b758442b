"""The oracle against the reference Fortran's own outputs (CPU).

The fixtures in tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference modules (mod_mesh, mod_advect3d_eq,
mod_dg_optr_kernel[_opt1]) compiled from their sources, on the inputs of
dgk.inputs.  The oracle is required to reproduce them bit for bit: it
evaluates every expression in the reference's order with no FMA contraction.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from dgk import inputs

GOLDEN = Path(__file__).resolve().parent / "golden"


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("p", range(1, 16))
def test_tendency_matches_reference(p):
    z = np.load(GOLDEN / f"tend_p{p}.npz")
    nx, ny, nz, pp = (int(v) for v in z["mesh"])
    assert pp == p
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=int(z["seed"]))
    assert digest(q, u, v, w) == str(z["input_sha256"])
    m = oracle.Mesh(nx, ny, nz, p)
    Q, U, V, W = m.field(q), m.field(u), m.field(v), m.field(w)
    assert np.array_equal(m.cal_tend(Q, U, V, W, oracle.GENERAL), z["general"])
    if p >= 8:  # OPT1 splits the l-sum into three partial sums above p = 7
        assert np.array_equal(m.cal_tend(Q, U, V, W, oracle.OPT1), z["opt1"])
    else:
        assert np.array_equal(m.cal_tend(Q, U, V, W, oracle.OPT1), z["general"])


def test_c1_time_loop_matches_reference():
    z = np.load(GOLDEN / "c1_run.npz")
    wl = inputs.c1_sine_noise()
    assert digest(wl.q, wl.u, wl.v, wl.w) == str(z["input_sha256"])
    assert wl.dt == float(z["dt"]) and wl.nstep == int(z["nstep"])
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    for var, key in ((oracle.GENERAL, "general"), (oracle.OPT1, "opt1")):
        q, mm, _ = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, wl.nstep, variant=var, trace=True)
        assert np.array_equal(q, z[f"{key}_q"])
        assert np.array_equal(mm, z[f"{key}_minmax"])


@pytest.mark.parametrize("name,make", [("c2", inputs.c2_gaussian), ("c3", inputs.c3_sine)])
def test_large_fixture_inputs_reproduce(name, make):
    """The 32^3 p=7 and 16^3 p=15 fixtures were written from these inputs."""
    z = np.load(GOLDEN / f"{name}_run.npz")
    wl = make()
    assert digest(wl.q, wl.u, wl.v, wl.w) == str(z["input_sha256"])
    assert wl.dt == float(z["dt"]) and wl.nstep == int(z["nstep"])
    assert np.array_equal(np.array(wl.vel), z["vel"])
