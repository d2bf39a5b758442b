"""Serving (dgk_serve_*): requests admitted by device footprint into one HBM
arena and advanced one SSP-RK3 step per tick beside the others.

* each served field equals the same solve run alone by a dgk_solver of its
  own form, bit for bit, whatever else was admitted, and under a budget that
  forces requests to wait for memory;
* the 1e-5 class stays within 1e-5 relative L2 of the FP64 solve;
* small requests agree with the CPU oracle to 1e-12;
* the device summary (min, max, quadrature mass and norm) equals the same
  quantities of the returned field, and the mass is conserved;
* a request that arrives while a long one runs is admitted and finishes
  first (continuous batching, not a queue);
* over-budget and malformed requests are refused with their status codes.
"""
import time

import numpy as np
import pytest
from numpy.polynomial import legendre

import dgk
import oracle
from dgk import inputs, serving

pytestmark = pytest.mark.gpu


def small_requests(n, seed, nstep_range=(3, 12), p_range=(2, 7), ne_range=(2, 5)):
    return inputs.serve_requests(n=n, seed=seed, rate=1000.0, ne_range=ne_range,
                                 p_range=p_range, nstep_range=nstep_range)


def lgl_weights(x1D):
    """w_i = 2 / (p (p+1) P_p(x_i)^2) on the nodes of operator_data/p{p}.dat."""
    p = len(x1D) - 1
    c = np.zeros(p + 1)
    c[-1] = 1.0
    return 2.0 / (p * (p + 1) * legendre.legval(x1D, c) ** 2)


def nodes(r):
    return int(np.prod(r["ne"])) * (r["p"] + 1) ** 3


def serve_all(reqs, **kw):
    outs = [np.empty(nodes(r)) for r in reqs]
    with serving.Server(**kw) as sv:
        for k, r in enumerate(reqs):
            sv.submit(r, q_out=outs[k], tag=k)
        res = sv.wait(len(reqs), timeout=300)
        st = sv.stats()
    by_tag = {d["tag"]: d for d in res}
    assert all(d["status"] == 0 for d in res), [d["message"] for d in res if d["status"]]
    return outs, [by_tag[k] for k in range(len(reqs))], st


@pytest.mark.parametrize("policy,max_steps", [(serving.TIME_FAIR, 16), (serving.STEP_FAIR, 1),
                                              (serving.TIME_FAIR, 3)])
def test_served_results_equal_solves_alone(policy, max_steps):
    reqs = small_requests(12, seed=21)
    outs, res, st = serve_all(reqs, hbm_budget=1 << 30, max_active=5, nstreams=3,
                              policy=policy, max_steps_per_tick=max_steps)
    assert st["completed"] == len(reqs) and st["active_peak"] <= 5
    for k, r in enumerate(reqs):
        alone = serving.solve_alone(r, res[k]["form"], res[k]["dt"])
        assert np.array_equal(outs[k], alone), k
        if res[k]["form"] == "HIP_MP":
            ref = serving.solve_alone(r, "HIP", res[k]["dt"])
            assert np.linalg.norm(outs[k] - ref) <= 1e-5 * np.linalg.norm(ref), k


def test_memory_bound_admission_keeps_results():
    reqs = small_requests(10, seed=22, ne_range=(3, 6))
    sizes = [serving.footprint(r["ne"], r["p"], r["tol_class"]) for r in reqs]
    budget = 2 * max(sizes)
    outs, res, st = serve_all(reqs, hbm_budget=budget, max_active=16, nstreams=2)
    assert st["bytes_peak"] <= budget
    assert st["active_peak"] < len(reqs)  # memory, not slots, limited admission
    for k, r in enumerate(reqs):
        assert res[k]["bytes"] == sizes[k]
        assert np.array_equal(outs[k], serving.solve_alone(r, res[k]["form"], res[k]["dt"])), k


def test_small_requests_against_oracle():
    reqs = [dict(r, tol_class=0) for r in small_requests(4, seed=23, ne_range=(2, 3))]
    outs, res, _ = serve_all(reqs, hbm_budget=256 << 20, max_active=4, nstreams=2)
    for k, r in enumerate(reqs):
        x1D = inputs.lgl_nodes(r["p"])
        X, Y, Z = inputs.node_coords(*r["ne"], x1D)
        q0 = inputs.gaussian_blob(X, Y, Z, r["centre"], r["sigma"])
        vel = [np.full_like(q0, c) for c in r["vel"]]
        m = oracle.Mesh(*r["ne"], r["p"])
        ref, _, _ = m.run(q0, *vel, res[k]["dt"], r["nstep"])
        ref = ref.reshape(-1)
        # the device blob differs from numpy's exp by rounding only
        assert np.abs(outs[k] - ref).max() <= 1e-12 * np.abs(ref).max(), k
        assert res[k]["dt"] == pytest.approx(inputs.cfl_dt(x1D, r["ne"], r["vel"]), rel=1e-15)


def test_summary_and_mass_conservation():
    reqs = small_requests(6, seed=24)
    reqs.append(dict(reqs[0], nstep=0))
    outs, res, _ = serve_all(reqs, hbm_budget=1 << 30, max_active=8, nstreams=2)
    for k, r in enumerate(reqs):
        x1D = inputs.lgl_nodes(r["p"])
        w = lgl_weights(x1D)
        Nq = r["p"] + 1
        w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).reshape(-1)
        jac = np.prod([1.0 / n for n in r["ne"]]) / 8.0
        q = outs[k].reshape(-1, Nq ** 3)
        assert res[k]["qmin"] == q.min() and res[k]["qmax"] == q.max()
        assert res[k]["mass"] == pytest.approx(jac * (q * w3).sum(), rel=1e-13)
        assert res[k]["l2sq"] == pytest.approx(jac * (q * q * w3).sum(), rel=1e-13)
        X, Y, Z = inputs.node_coords(*r["ne"], x1D)
        q0 = inputs.gaussian_blob(X, Y, Z, r["centre"], r["sigma"]).reshape(-1, Nq ** 3)
        mass0 = jac * (q0 * w3).sum()
        tol = 1e-12 if res[k]["form"] == "HIP" else 1e-6
        assert abs(res[k]["mass"] - mass0) <= tol * abs(mass0), k


def test_late_arrival_joins_running_mix():
    long_req = dict(ne=(16, 16, 16), p=7, nstep=1000, tol_class=0, vel=(1.0, 0.5, 0.25),
                    centre=(0.5, 0.5, 0.5), sigma=0.1)
    short = small_requests(4, seed=25, nstep_range=(5, 10))
    with serving.Server(hbm_budget=4 << 30, max_active=8, nstreams=4) as sv:
        sv.submit(long_req, tag=100)
        t_end = time.monotonic() + 60
        while sv.stats()["admitted"] < 1 and time.monotonic() < t_end:
            time.sleep(0.001)
        for k, r in enumerate(short):
            sv.submit(r, tag=k)
        res = {d["tag"]: d for d in sv.wait(5, timeout=300)}
    assert all(d["status"] == 0 for d in res.values())
    for k in range(len(short)):
        assert res[k]["t_admit"] < res[100]["t_done"]
        assert res[k]["t_done"] < res[100]["t_done"], k


def test_refusals():
    with serving.Server(hbm_budget=8 << 20, max_active=2, nstreams=1) as sv:
        big = dict(ne=(16, 16, 16), p=7, nstep=1, vel=(1, 0, 0), centre=(0.5,) * 3, sigma=0.1)
        with pytest.raises(dgk.DgkError) as ei:
            sv.submit(big)
        assert ei.value.code == dgk.ERR_NOMEM
        ok = dict(ne=(2, 2, 2), p=2, nstep=2, vel=(1, 0, 0), centre=(0.5,) * 3, sigma=0.1)
        for bad, code in ((dict(ok, p=20), dgk.ERR_UNSUPPORTED), (dict(ok, sigma=0.0), dgk.ERR_ARG),
                          (dict(ok, ne=(0, 2, 2)), dgk.ERR_ARG),
                          (dict(ok, vel=(0, 0, 0)), dgk.ERR_ARG),
                          (dict(ok, tol_class=7), dgk.ERR_ARG)):
            with pytest.raises(dgk.DgkError) as ei:
                sv.submit(bad)
            assert ei.value.code == code, bad
        sv.submit(ok)
        (d,) = sv.wait(1, timeout=60)
        assert d["status"] == 0
