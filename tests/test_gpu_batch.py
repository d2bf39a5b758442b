"""Batched ensembles (dgk_batch_*): many independent boxes of mixed size and
order advanced together, one stage launch per order group, one RK step
captured in a hipGraph and replayed.

* every instance equals the same box run alone by a dgk_solver, bit for bit
  (the per-element arithmetic does not depend on the batch's composition);
* graph replays equal eager launches after every replay, with the host's own
  reads and synchronisations in between;
* the device initial condition equals the numpy generator (to exp rounding);
* HIP_MP instances stay within 1e-5 relative L2 of the FP64 ones;
* small instances agree with the CPU oracle to 1e-12.
"""
import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

pytestmark = pytest.mark.gpu


def small_ensemble(n=10, seed=11):
    return inputs.ensemble(n=n, seed=seed, ne_choices=(2, 3, 4), p_choices=(2, 3, 5, 7),
                           nstep=6)


def make_batch(form, ens):
    b = dgk.Batch(form, ens)
    for k, it in enumerate(ens):
        b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
    return b


def test_batch_equals_instances_alone():
    ens = small_ensemble()
    with make_batch("HIP", ens) as b:
        b.step(ens[0]["nstep"])
        got = [b.get_field(k) for k in range(len(ens))]
    for k, it in enumerate(ens):
        ne, p = it["ne"], it["p"]
        with dgk.Solver.structured("HIP", *ne, p) as s:
            s.init_gaussian(it["centre"], it["sigma"], it["vel"])
            s.step(it["dt"], it["nstep"])
            alone = s.get_field(dgk.FIELD_Q)
        assert np.array_equal(got[k], alone), k


def test_graph_replays_equal_eager_launches():
    ens = small_ensemble(n=8, seed=12)
    with make_batch("HIP", ens) as g, make_batch("HIP", ens) as e:
        e.set_graph(False)
        for _ in range(3):
            g.step(2)
            e.step(2)
            for k in range(len(ens)):
                assert np.array_equal(g.get_field(k), e.get_field(k)), k


def test_device_initial_condition_matches_generator():
    ens = small_ensemble(n=4, seed=13)
    with make_batch("HIP", ens) as b:
        for k, it in enumerate(ens):
            x1D = inputs.lgl_nodes(it["p"])
            X, Y, Z = inputs.node_coords(*it["ne"], x1D)
            ref = inputs.gaussian_blob(X, Y, Z, it["centre"], it["sigma"])
            got = b.get_field(k)
            assert np.abs(got - ref).max() <= 4e-16
            for c, which in enumerate((dgk.FIELD_U, dgk.FIELD_V, dgk.FIELD_W)):
                assert np.all(b.get_field(k, which) == it["vel"][c])


def test_mixed_precision_batch_within_tolerance():
    ens = small_ensemble(n=10, seed=14)
    with make_batch("HIP", ens) as a, make_batch("HIP_MP", ens) as m:
        a.step(6)
        m.step(6)
        for k in range(len(ens)):
            ref, got = a.get_field(k), m.get_field(k)
            assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref), k


def test_small_instances_against_oracle():
    ens = small_ensemble(n=6, seed=15)
    with make_batch("HIP", ens) as b:
        init = [(b.get_field(k), b.get_field(k, dgk.FIELD_U), b.get_field(k, dgk.FIELD_V),
                 b.get_field(k, dgk.FIELD_W)) for k in range(len(ens))]
        b.step(ens[0]["nstep"])
        for k, it in enumerate(ens):
            m = oracle.Mesh(*it["ne"], it["p"])
            ref, _, _ = m.run(*init[k], it["dt"], it["nstep"])
            got = b.get_field(k)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), k


@pytest.mark.parametrize("form", ["HIP", "HIP_MP"])
def test_constant_and_field_velocities(form):
    """Boxes set up by init_gaussian have one constant velocity, which their
    order group reads from the geometry records.  A varying velocity field
    given to one instance puts its group back on the fields; a constant
    field given afterwards puts it back on the records.  Every instance
    equals its solve alone, bit for bit, in each state."""
    ens = small_ensemble(n=6, seed=16)
    it0 = ens[0]
    n0 = int(np.prod(it0["ne"])) * (it0["p"] + 1) ** 3
    ramp = it0["vel"][0] * (1.0 + 0.1 * inputs.u01(77, np.arange(n0, dtype=np.uint64)))
    for u0 in (ramp, np.full(n0, 0.25)):
        with make_batch(form, ens) as b:
            b.set_field(0, dgk.FIELD_U, u0)
            b.step(4)
            got = [b.get_field(k) for k in range(len(ens))]
        for k, it in enumerate(ens):
            with dgk.Solver.structured(form, *it["ne"], it["p"]) as s:
                s.init_gaussian(it["centre"], it["sigma"], it["vel"])
                if k == 0:
                    s.set_field(dgk.FIELD_U, u0)
                s.step(it["dt"], 4)
                alone = s.get_field(dgk.FIELD_Q)
            assert np.array_equal(got[k], alone), (k, u0[0] == u0[-1])


@pytest.mark.parametrize("form", ["HIP", "HIP_MP"])
def test_autotune_skips_inapplicable_candidates(form):
    # batched boxes with constant velocities cannot take the kernel over runs
    # of elements (a candidate at low order): the autotuner does not time it
    # (0) on the batch, does on a plain solver, and never reports it for the
    # batch
    p = 3
    last = len(dgk.tile_shapes(form, p)) - 1
    run = [c for c in range(last + 1) if dgk.kernel_name(form, p, c).startswith("stage_run")]
    assert len(run) == 1
    run = run[0]
    ens = inputs.ensemble(n=4, seed=12, ne_choices=(4,), p_choices=(p,), nstep=2)
    with dgk.Batch(form, ens) as b:
        for k, it in enumerate(ens):
            b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        b.step(1)
        rec = b.autotune(reps=2)
        assert [x["p"] for x in rec] == [p]
        assert rec[0]["candidate"] != run and rec[0]["us_per_launch"] > 0
        assert rec[0]["us_default"] > 0
    wl = inputs.c1_sine_noise()
    with dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        r = s.autotune(reps=2)
        assert len(r["us_per_launch"]) == last + 1 and all(t > 0 for t in r["us_per_launch"])
