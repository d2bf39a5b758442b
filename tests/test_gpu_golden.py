"""The FP64 device path over the full C2 and C3 runs.

* tests/golden/c2_run.npz and c3_run.npz were written by the reference
  modules (GENERAL and OPT1) on 32^3 elements at p=7 (200 steps) and 16^3 at
  p=15 (100 steps).  The device result must agree within max-norm relative
  error 1e-12 on the stored sample, the L2 norm and the min/max trace.
* The whole final field (16.8M DOF) must agree element-wise within 1e-12
  max-norm relative with the CPU oracle's run of the same loop (main.f90:51-80)
  on the same inputs: GENERAL at both sizes, and OPT1's p>7 summation order
  (mod_dg_optr_kernel_opt1.F90.erb:242-274) at p=15, where it rounds
  differently.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
TOL = 1e-12
CASES = [("c2", inputs.c2_gaussian, ("general",)), ("c3", inputs.c3_sine, ("general", "opt1"))]


@pytest.mark.parametrize("name,make,variants", CASES, ids=[c[0] for c in CASES])
def test_full_run_matches_reference(name, make, variants):
    z = np.load(GOLDEN / f"{name}_run.npz")
    wl = make()
    print(f"{name}: {wl.NeX}^3 p={wl.p} {wl.nstep} steps on the device", file=sys.stderr,
          flush=True)
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
    # the whole field against the oracle's full run
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    for key in variants:
        variant = oracle.GENERAL if key == "general" else oracle.OPT1
        print(f"{name}: oracle {key} run on {oracle.num_threads()} threads", file=sys.stderr,
              flush=True)
        ref, ref_mm, secs = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, wl.nstep, variant=variant,
                                  trace=True)
        ref = ref.reshape(-1)
        err = np.abs(q - ref).max() / np.abs(ref).max()
        print(f"{name}: whole-field max rel err vs oracle {key} {err:.3e} ({secs:.1f} s)",
              file=sys.stderr, flush=True)
        assert err <= TOL, (key, err)
        assert np.abs(mm - ref_mm).max() <= TOL * np.abs(ref_mm).max(), key
