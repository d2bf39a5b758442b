"""The FP64 device path against the reference Fortran's own full-size runs.

tests/golden/c2_run.npz and c3_run.npz were written by the reference modules
(GENERAL and OPT1) on 32^3 elements at p=7 (200 steps) and 16^3 at p=15
(100 steps).  The device result must agree within max-norm relative error
1e-12 on the stored sample, the L2 norm and the min/max trace.
"""
from pathlib import Path

import numpy as np
import pytest

import dgk
from dgk import inputs

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
TOL = 1e-12


@pytest.mark.parametrize("name,make", [("c2", inputs.c2_gaussian), ("c3", inputs.c3_sine)])
def test_full_run_matches_reference(name, make):
    z = np.load(GOLDEN / f"{name}_run.npz")
    wl = make()
    with dgk.Solver.structured("HIP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        mm = []
        for _ in range(wl.nstep):
            s.step(wl.dt, 1)
            mm.append(s.minmax())
        q = s.get_field(dgk.FIELD_Q).reshape(-1)
    mm = np.array(mm)
    for key in ("general", "opt1"):
        ref = z[f"{key}_sample"]
        got = q[z[f"{key}_idx"]]
        assert np.abs(got - ref).max() <= TOL * np.abs(ref).max(), key
        assert abs(np.sqrt(np.sum(q * q)) - float(z[f"{key}_l2"])) <= TOL * float(z[f"{key}_l2"])
        assert np.abs(mm - z[f"{key}_minmax"]).max() <= TOL * np.abs(z[f"{key}_minmax"]).max()
