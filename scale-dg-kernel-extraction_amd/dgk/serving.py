"""Serving: continuous batching of advection solves (dgk_serve_*).

`Server` wraps the C ABI: requests are periodic boxes with a Gaussian-blob
initial condition and constant velocity, admitted by device footprint into a
fixed HBM arena and advanced one SSP-RK3 step per tick beside whatever else is
resident.  `run_stream` replays a seeded Poisson stream (inputs.serve_requests)
against a server in real time and reports latency percentiles, throughput and
the error of a sample of results against the same solves run alone.
"""
from __future__ import annotations

import ctypes as C
import sys
import time

import numpy as np

from . import (EVAL_HIP, EVAL_HIP_MP, FIELD_Q, OPERATOR_DIR, Solver, _check, _ptr, lib)
from . import inputs

TOL_FP64, TOL_1E5 = 0, 1
STEP_FAIR, TIME_FAIR = 0, 1


class Request(C.Structure):
    _fields_ = [("ne", C.c_int32 * 3), ("p", C.c_int32), ("nstep", C.c_int32),
                ("tol_class", C.c_int32), ("dom", C.c_double * 3), ("dt", C.c_double),
                ("vel", C.c_double * 3), ("centre", C.c_double * 3), ("sigma", C.c_double),
                ("q_out", C.c_void_p), ("tag", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("tag", C.c_int64), ("status", C.c_int32),
                ("evaltype", C.c_int32), ("p", C.c_int32), ("nstep", C.c_int32),
                ("dof", C.c_int64), ("bytes", C.c_int64), ("dt", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("qmin", C.c_double), ("qmax", C.c_double), ("mass", C.c_double),
                ("l2sq", C.c_double), ("message", C.c_char * 160)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "submitted", "admitted", "completed", "failed", "queued", "active", "active_peak",
        "ticks", "bytes_in_use", "bytes_peak", "budget", "bypasses")] + [
        ("dof_updates", C.c_double)]


def footprint(ne, p, tol_class=TOL_FP64) -> int:
    """Device bytes a request of this shape holds while admitted."""
    a = (C.c_int32 * 3)(*ne)
    b = C.c_int64(0)
    _check(lib().dgk_serve_footprint(a, int(p), int(tol_class), C.byref(b)))
    return b.value


def _result_dict(r: Result) -> dict:
    d = {name: getattr(r, name) for name, _ in Result._fields_ if name != "message"}
    d["message"] = r.message.decode(errors="replace")
    d["form"] = "HIP_MP" if r.evaltype == EVAL_HIP_MP else "HIP"
    return d


class Server:
    """A serving loop on the current device.

    hbm_budget: bytes of the device arena (allocated once at creation);
    max_active: requests resident at once; nstreams: HIP streams."""

    def __init__(self, hbm_budget: int = 16 << 30, max_active: int = 64, nstreams: int = 4,
                 policy: int = TIME_FAIR, max_steps_per_tick: int = 64,
                 operator_dir=OPERATOR_DIR):
        h = C.c_void_p()
        _check(lib().dgk_serve_create(str(operator_dir).encode(), int(hbm_budget),
                                      int(max_active), int(nstreams), C.byref(h)))
        self._h = h
        self._outputs = {}  # id -> host array receiving the final q
        self.set_policy(policy, max_steps_per_tick)

    def set_policy(self, policy: int, max_steps_per_tick: int = 64) -> None:
        _check(lib().dgk_serve_set_policy(self._h, int(policy), int(max_steps_per_tick)))

    def submit(self, req: dict, q_out: np.ndarray | None = None, tag: int = 0) -> int:
        """Queue one request (keys ne, p, nstep, tol_class, vel, centre, sigma,
        optional dom and dt).  q_out, if given, is a float64 array of Ne*Np
        values that receives the final field."""
        r = Request()
        r.ne[:] = [int(x) for x in req["ne"]]
        r.p = int(req["p"])
        r.nstep = int(req["nstep"])
        r.tol_class = int(req.get("tol_class", TOL_FP64))
        r.dom[:] = [float(x) for x in req.get("dom", (1.0, 1.0, 1.0))]
        r.dt = float(req.get("dt", 0.0))
        r.vel[:] = [float(x) for x in req["vel"]]
        r.centre[:] = [float(x) for x in req["centre"]]
        r.sigma = float(req["sigma"])
        r.tag = int(tag)
        if q_out is not None:
            n = int(np.prod(req["ne"])) * (int(req["p"]) + 1) ** 3
            assert q_out.dtype == np.float64 and q_out.flags.c_contiguous and q_out.size == n
            r.q_out = q_out.ctypes.data
        i = C.c_int64(0)
        _check(lib().dgk_serve_submit(self._h, C.byref(r), C.byref(i)))
        if q_out is not None:
            self._outputs[i.value] = q_out
        return i.value

    def poll(self, timeout: float = 0.0, max_results: int = 256) -> list:
        buf = (Result * max_results)()
        n = C.c_int(0)
        _check(lib().dgk_serve_poll(self._h, buf, max_results, C.byref(n), float(timeout)))
        out = []
        for k in range(n.value):
            d = _result_dict(buf[k])
            d["q"] = self._outputs.pop(d["id"], None)
            out.append(d)
        return out

    def wait(self, count: int, timeout: float = 600.0) -> list:
        """Poll until `count` results have arrived (raises on timeout)."""
        got, t_end = [], time.monotonic() + timeout
        while len(got) < count:
            left = t_end - time.monotonic()
            if left <= 0:
                raise TimeoutError(f"{len(got)} of {count} results after {timeout} s")
            got += self.poll(timeout=min(left, 1.0))
        return got

    def stats(self) -> dict:
        s = Stats()
        _check(lib().dgk_serve_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in Stats._fields_}

    def close(self) -> None:
        if self._h:
            _check(lib().dgk_serve_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_alone(req: dict, form: str, dt: float) -> np.ndarray:
    """The same solve on a dgk_solver of its own (plain path, no server)."""
    with Solver.structured(form, *req["ne"], req["p"]) as s:
        s.init_gaussian(req["centre"], req["sigma"], req["vel"])
        s.step(dt, req["nstep"])
        return s.get_field(FIELD_Q).reshape(-1)


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def run_stream(n: int = 200, rate: float = 4.0, seed: int = 5, hbm_budget: int = 64 << 30,
               max_active: int = 64, nstreams: int = 4, sample: int | None = None,
               sample_max_dof: int | None = None, progress: bool = True,
               policy: int = TIME_FAIR, max_steps_per_tick: int = 64,
               return_fields: bool = False) -> dict:
    """Replay the first n requests of the seeded Poisson stream in real time.

    Every request (sample=None), or up to `sample` requests of each tolerance
    class with at most sample_max_dof DOF, returns its final field; after the
    stream each of them is solved again alone (dgk_solver) in its own form and,
    for the 1e-5 class, in plain FP64, to show that the served result equals
    the solve alone and to state its error against the FP64 result.  With
    return_fields the served fields and the FP64 solves are returned too
    (out["fields"][k] = (served, fp64))."""
    reqs = inputs.serve_requests(n=n, seed=seed, rate=rate)
    picked = {TOL_FP64: [], TOL_1E5: []}
    for k, r in enumerate(reqs):
        dof = int(np.prod(r["ne"])) * (r["p"] + 1) ** 3
        if sample_max_dof is not None and dof > sample_max_dof:
            continue
        if sample is None or len(picked[r["tol_class"]]) < sample:
            picked[r["tol_class"]].append(k)
    wanted = set(picked[TOL_FP64] + picked[TOL_1E5])
    results = {}
    with Server(hbm_budget=hbm_budget, max_active=max_active, nstreams=nstreams,
                policy=policy, max_steps_per_tick=max_steps_per_tick) as sv:
        t0 = time.perf_counter()
        nxt, last_note = 0, t0
        while len(results) < n:
            now = time.perf_counter() - t0
            while nxt < n and reqs[nxt]["arrival"] <= now:
                r = reqs[nxt]
                q = None
                if nxt in wanted:
                    q = np.empty(int(np.prod(r["ne"])) * (r["p"] + 1) ** 3)
                sv.submit(r, q_out=q, tag=nxt)
                nxt += 1
            wait = (reqs[nxt]["arrival"] - now) if nxt < n else 1.0
            for d in sv.poll(timeout=max(0.0, min(wait, 1.0))):
                results[d["tag"]] = d
            if progress and time.perf_counter() - last_note > 10.0:
                last_note = time.perf_counter()
                st = sv.stats()
                print(f"serve: {len(results)}/{n} done, {st['active']} active, "
                      f"{st['queued']} queued, {last_note - t0:.0f} s", file=sys.stderr,
                      flush=True)
        stats = sv.stats()
    failed = [d for d in results.values() if d["status"] != 0]
    if failed:
        raise RuntimeError(f"{len(failed)} requests failed, first: {failed[0]['message']}")
    done = [results[k] for k in range(n)]
    lat = [d["t_done"] - d["t_submit"] for d in done]
    queue = [d["t_admit"] - d["t_submit"] for d in done]
    t_first = min(d["t_submit"] for d in done)
    t_last = max(d["t_done"] for d in done)
    work = sum(d["dof"] * d["nstep"] for d in done)
    out = {
        "requests": n, "rate_per_s": rate, "seed": seed,
        "p50_latency_s": _pct(lat, 50), "p99_latency_s": _pct(lat, 99),
        "mean_latency_s": float(np.mean(lat)), "max_latency_s": float(np.max(lat)),
        "p50_queue_s": _pct(queue, 50), "p99_queue_s": _pct(queue, 99),
        "DOF-updates/s": work / (t_last - t_first), "makespan_s": t_last - t_first,
        "dof_updates": float(work),
        "by_class": {},
        "server": {k: stats[k] for k in ("active_peak", "bytes_peak", "budget", "ticks",
                                          "bypasses")},
        "hbm_budget": hbm_budget, "max_active": max_active, "nstreams": nstreams,
        "policy": "time-fair" if policy == TIME_FAIR else "step-fair",
        "max_steps_per_tick": max_steps_per_tick,
    }
    for cls, name in ((TOL_FP64, "fp64"), (TOL_1E5, "1e-5")):
        sel = [lat[k] for k in range(n) if reqs[k]["tol_class"] == cls]
        out["by_class"][name] = {"count": len(sel), "p50_latency_s": _pct(sel, 50),
                                 "p99_latency_s": _pct(sel, 99)}
    # served results against the same solves alone
    errs, fields = [], {}
    for k in sorted(wanted):
        d, r = results[k], reqs[k]
        alone = solve_alone(r, d["form"], d["dt"])
        ref = alone if d["form"] == "HIP" else solve_alone(r, "HIP", d["dt"])
        q = d["q"]
        if return_fields:
            fields[k] = (q, ref)
        errs.append({
            "request": k, "form": d["form"], "dof": d["dof"], "nstep": d["nstep"],
            "same_as_alone": bool(np.array_equal(q, alone)),
            "rel_l2_vs_fp64": float(np.linalg.norm(q - ref) / np.linalg.norm(ref)),
            "max_rel_vs_fp64": float(np.abs(q - ref).max() / np.abs(ref).max())})
    out["sample_vs_alone"] = {
        "checked": len(errs),
        "all_same_as_alone": all(e["same_as_alone"] for e in errs),
        "max_rel_l2_1e-5_class": max([e["rel_l2_vs_fp64"] for e in errs
                                      if e["form"] == "HIP_MP"], default=None),
        "max_rel_fp64_class": max([e["max_rel_vs_fp64"] for e in errs
                                   if e["form"] == "HIP"], default=None),
        "requests": errs}
    if return_fields:
        out["fields"] = fields
        out["reqs"] = reqs
        out["results"] = done
    return out
