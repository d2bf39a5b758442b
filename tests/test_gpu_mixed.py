"""The mixed-precision device path ("HIP_MP": FP32 MFMA contractions and Lift
with FP32 accumulation, FP64 state / flux jumps / RK) against the FP64
references.

Tolerance (stated by the workload definition): relative L2 error
||q_mp - q_ref||_2 / ||q_ref||_2 <= 1e-5 after the time loop; a single
tendency evaluation is held to the same bound relative to ||dqdt_ref||_2.
"""
from pathlib import Path

import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("p", range(1, 16))
def test_mp_tendency_every_order(p):
    nx, ny, nz = (3, 2, 2) if p <= 6 else (2, 2, 2) if p <= 11 else (2, 1, 2)
    m = oracle.Mesh(nx, ny, nz, p)
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=300 + p)
    ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    with dgk.Solver.create("HIP_MP", m.Nq, m.Ne, m.NeA, m.NhaloNode, m.D, m.Lift, m.VMapM,
                           m.VMapP, m.halo_src_map, m.normal_fn, m.Escale, m.Fscale) as s:
        s.set_fields(q, u, v, w)
        got = s.cal_tend()
    assert rel_l2(got, ref) <= TOL


def test_mp_c1_run_against_oracle():
    wl = inputs.c1_sine_noise()
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    ref, _, _ = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, wl.nstep)
    with dgk.Solver.structured("HIP_MP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.step(wl.dt, wl.nstep)
        got = s.get_field(dgk.FIELD_Q)
    assert rel_l2(got, ref) <= TOL


def test_mp_c3_p15_against_reference_fixture():
    """16^3 elements at p=15, 100 steps: against the reference's own run."""
    z = np.load(GOLDEN / "c3_run.npz")
    wl = inputs.c3_sine()
    with dgk.Solver.structured("HIP_MP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.step(wl.dt, wl.nstep)
        q = s.get_field(dgk.FIELD_Q).reshape(-1)
    ref = z["general_sample"]
    got = q[z["general_idx"]]
    assert rel_l2(got, ref) <= TOL
    assert abs(np.sqrt(np.sum(q * q)) - float(z["general_l2"])) <= TOL * float(z["general_l2"])


def test_mp_c2_p7_against_fp64_device_path():
    wl = inputs.c2_gaussian()
    out = {}
    for form in ("HIP", "HIP_MP"):
        with dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
            s.set_fields(wl.q, wl.u, wl.v, wl.w)
            s.step(wl.dt, wl.nstep)
            out[form] = s.get_field(dgk.FIELD_Q)
    assert rel_l2(out["HIP_MP"], out["HIP"]) <= TOL


def test_mp_rejects_non_separable_lift():
    m = oracle.Mesh(2, 2, 2, 3)
    L = m.Lift.copy()
    L[0, 1, 2, 3] += 0.5
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("HIP_MP", m.Nq, m.Ne, m.NeA, m.NhaloNode, m.D, L, m.VMapM, m.VMapP,
                          m.halo_src_map, m.normal_fn, m.Escale, m.Fscale)
    assert ei.value.code == dgk.ERR_UNSUPPORTED


def test_fp64_stored_lift_kernel_for_non_separable_lift():
    """A Lift_mat that is not of tensor-product form runs through the
    stored-Lift FP64 kernel and still matches the oracle."""
    m = oracle.Mesh(2, 3, 2, 4)
    L = m.Lift.copy()
    L[2] += 0.01 * np.random.default_rng(3).standard_normal(L[2].shape)
    m.Lift = np.ascontiguousarray(L)
    q, u, v, w = inputs.random_fields(2, 3, 2, 4, seed=77)
    ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    with dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, m.D, m.Lift, m.VMapM,
                           m.VMapP, m.halo_src_map, m.normal_fn, m.Escale, m.Fscale) as s:
        s.set_fields(q, u, v, w)
        got = s.cal_tend()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("p", range(1, 20))
def test_mp_tile_shapes_give_identical_results(p):
    """Every tile shape (elements per workgroup) gives the same bits, also with
    a partial last workgroup (30 elements)."""
    ne = (3, 2, 5) if p <= 7 else (2, 1, 3)
    q, u, v, w = inputs.random_fields(*ne, p, seed=700 + p, vel_scale=0.8)
    dt = 0.2 * inputs.cfl_dt(inputs.lgl_nodes(p), ne, (0.8, 0.8, 0.8))
    got = []
    try:
        for cand in range(len(dgk.tile_shapes("HIP_MP", p))):
            dgk.tile_select("HIP_MP", p, cand)
            with dgk.Solver.structured("HIP_MP", *ne, p) as s:
                s.set_fields(q, u, v, w)
                s.step(dt, 3)
                got.append(s.get_field(dgk.FIELD_Q))
    finally:
        dgk.tile_select("HIP_MP", p, -1)
    for k in range(1, len(got)):
        assert np.array_equal(got[k], got[0]), k


def test_autotune_solver_and_batch():
    wl = inputs.c1_sine_noise()
    with dgk.Solver.structured("HIP_MP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.step(wl.dt, 2)
        before = s.get_field(dgk.FIELD_Q)
        default = dgk.tile_selected("HIP_MP", wl.p)
        r = s.autotune(reps=3)
        assert 0 <= r["best"] < len(dgk.tile_shapes("HIP_MP", wl.p))
        assert len(r["us_per_launch"]) == len(dgk.tile_shapes("HIP_MP", wl.p))
        assert all(t > 0 for t in r["us_per_launch"])
        assert dgk.tile_selected("HIP_MP", wl.p) == default  # the choice is the solver's own
        assert np.array_equal(s.get_field(dgk.FIELD_Q), before)  # state untouched
        s.step(wl.dt, 2)
        tuned = s.get_field(dgk.FIELD_Q)
    with dgk.Solver.structured("HIP_MP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.step(wl.dt, 4)
        assert np.array_equal(s.get_field(dgk.FIELD_Q), tuned)  # shapes give the same bits
    ens = inputs.ensemble(n=6, seed=31, ne_choices=(2, 3), p_choices=(3, 7), nstep=4)
    with dgk.Batch("HIP_MP", ens) as b, dgk.Batch("HIP_MP", ens) as ref:
        for bb in (b, ref):
            for k, it in enumerate(ens):
                bb.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        ref.step(2)
        b.step(1)
        rec = b.autotune(reps=3)
        assert sorted(x["p"] for x in rec) == sorted({it["p"] for it in ens})
        for x in rec:
            assert x["epb"] in dgk.tile_shapes("HIP_MP", x["p"]) and x["us_per_launch"] > 0
            # the default's time is the default candidate's own timing
            if x["candidate"] == dgk.tile_selected("HIP_MP", x["p"]):
                assert x["us_default"] == x["us_per_launch"]
            else:
                assert x["us_per_launch"] <= 0.98 * x["us_default"]
        b.step(1)  # re-captured with the tuned shapes
        for k in range(len(ens)):
            assert np.array_equal(b.get_field(k), ref.get_field(k)), k
