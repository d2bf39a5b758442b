#!/usr/bin/env python3
"""Benchmark of the DG advection hot path on one MI355X.

Headline workload (`value`): the plain FP64 path ("HIP") on the 32^3-element
periodic cube at p=7 (16,777,216 DOF), velocity U[-1,1]^3 and a Gaussian blob
from seed 2, dt from CFL 0.1.  A "step" is one SSP-RK3 time step of the whole
mesh (3 fused stage kernels).  Inputs are resident in HBM before the timed
region.  Metric: DOF-updates/s = Ne*Np*K / t for K timed steps.

The JSON line also carries
  roofline      the FP64 stage kernel: algorithmic HBM bytes per launch
                (136/3 B per DOF) over its average launch time, from HIP events
                recorded on the solver's stream around every launch; traffic
                from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/;
  cpu_baseline  the reference Fortran GENERAL path (oracle/_ref, OpenMP on the
                host cores) on a bounded number of steps of the same mesh;
  extra         the mixed-precision form on the same mesh; the batched
                ensemble (512 seeded instances, Ne/axis in {4,6,8,12,16},
                p in {3,5,7,11}, 50 steps, graph-replayed) in FP64 and mixed
                precision against the plain FP64 path run instance by
                instance; serving: the first --serve requests of the seeded
                Poisson stream (Ne/axis 4..32, p 2..15, nstep 10..500, half
                of them in the 1e-5 class) at --serve-rate requests/s, with
                p50/p99 latency, DOF-updates/s and a sample of results checked
                against the same solves alone, plus a burst of --serve-burst
                requests for the serving capacity.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "scale-dg-kernel-extraction_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_DOF_UPDATE = {"HIP": 136.0, "HIP_MP": 100.0}
METRIC = "DOF-updates/s plain FP64 vs batched mixed-prec; p50/p99 solve latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--form", default="HIP", choices=["HIP", "HIP_MP"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=120,
                    help="time steps of the bounded CPU baseline sample")
    ap.add_argument("--no-extra", action="store_true", help="headline only")
    ap.add_argument("--serve", type=int, default=150,
                    help="requests of the seeded Poisson stream to serve (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=6.0,
                    help="arrival rate of the serving stream, requests/s")
    ap.add_argument("--serve-burst", type=int, default=60,
                    help="requests submitted at once to measure serving capacity (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (the reference "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def copy_calibration(nbytes):
    """hipMemcpy device-to-device of a known size, five times."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 5}), flush=True)


def traffic_from_profile(form):
    """HBM bytes per launch measured by rocprofv3 PMC runs (profiles/)."""
    name = "stage_f64" if form == "HIP" else "stage_mp"
    f = ROOT / "profiles" / f"traffic_{name}_c2_32cube_p7.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def time_solver(dgk, inputs, form, steps, warmup):
    wl = inputs.c2_gaussian()
    dof = wl.Ne * wl.Np
    s = dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p)
    s.set_fields(wl.q, wl.u, wl.v, wl.w)
    s.step(wl.dt, warmup)
    s.synchronize()
    s.timer_reset()
    s.set_kernel_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(wl.dt, steps)
    s.synchronize()
    t1 = time.perf_counter()
    s.set_kernel_timing(False)
    tm = s.timers()
    qmin, qmax = s.minmax()
    s.close()
    secs = t1 - t0
    stage_s, stage_n = tm["Stage (fused)"]
    return wl, dict(value=dof * steps / secs, seconds=secs, launch_s=stage_s / max(stage_n, 1),
                    launches=stage_n, timers=tm, minmax=[qmin, qmax], dof=dof)


def batch_leg(dgk, inputs):
    """Batched ensemble vs the plain FP64 path run instance by instance."""
    ens = inputs.ensemble()
    out = {"instances": len(ens), "nstep": ens[0]["nstep"]}
    dof = np.array([(it["p"] + 1) ** 3 * it["ne"][0] ** 3 for it in ens], dtype=np.float64)
    nstep = ens[0]["nstep"]
    for form in ("HIP", "HIP_MP"):
        b = dgk.Batch(form, ens)
        for k, it in enumerate(ens):
            b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        out.setdefault("tile_shapes", {})[form] = b.autotune(reps=10)
        b.step(1)  # graph capture + warm-up
        b.synchronize()
        b.timer()
        t0 = time.perf_counter()
        b.step(nstep)
        b.synchronize()
        t1 = time.perf_counter()
        dev_s, _ = b.timer()
        out[form] = {"DOF-updates/s": float(dof.sum() * nstep / (t1 - t0)),
                     "seconds": t1 - t0, "device_seconds": dev_s,
                     "device_bytes": b.size()["device_bytes"]}
        b.close()
    # plain FP64 path, one solver per instance, on the p = 7 instances and a
    # sample of the rest (extrapolated by DOF)
    p7 = [k for k, it in enumerate(ens) if it["p"] == 7]
    solo_s = 0.0
    for k in p7:
        it = ens[k]
        s = dgk.Solver.structured("HIP", *it["ne"], it["p"])
        s.init_gaussian(it["centre"], it["sigma"], it["vel"])
        s.step(it["dt"], 1)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(it["dt"], nstep)
        s.synchronize()
        solo_s += time.perf_counter() - t0
        s.close()
    p7_dof = float(dof[p7].sum())
    out["plain_fp64_solo_p7"] = {"DOF-updates/s": p7_dof * nstep / solo_s, "seconds": solo_s,
                                 "instances": len(p7)}
    # batched forms on the p = 7 instances alone
    sub = [ens[k] for k in p7]
    for form in ("HIP", "HIP_MP"):
        b = dgk.Batch(form, sub)
        for k, it in enumerate(sub):
            b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        b.autotune(reps=10)
        b.step(1)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(nstep)
        b.synchronize()
        out[f"batched_{form}_p7"] = {"DOF-updates/s": p7_dof * nstep / (time.perf_counter() - t0)}
        b.close()
    out["speedup_batched_mp_vs_plain_fp64_p7"] = (
        out["batched_HIP_MP_p7"]["DOF-updates/s"] / out["plain_fp64_solo_p7"]["DOF-updates/s"])
    return out


def cpu_baseline(wl, steps):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    dof = wl.Ne * wl.Np
    sample = (f"{wl.NeX}^3 elements p={wl.p} ({dof} DOF), {steps} SSP-RK3 steps of the "
              f"same inputs (the GPU line times {wl.nstep}-step runs)")
    if oracle.reference_available():
        _, _, secs = oracle.run_reference("time", wl.NeX, wl.NeY, wl.NeZ, wl.p, "GENERAL",
                                          steps, wl.dt, wl.q, wl.u, wl.v, wl.w,
                                          threads=threads)
        kind = "reference"
    else:
        m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
        _, _, secs = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, steps)
        kind = "port"
    return {"value": dof * steps / secs, "unit": "DOF-updates/s", "cores": threads,
            "kind": kind, "sample": sample, "seconds": secs}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; a multi-GPU layout is "
                                   "described in DESIGN.md, not provided as a run mode"}))
        sys.exit(2)
    import dgk
    from dgk import inputs

    wl, r = time_solver(dgk, inputs, args.form, args.steps, args.warmup)
    dof = r["dof"]
    bytes_per_launch = BYTES_PER_DOF_UPDATE[args.form] / 3.0 * dof
    achieved = bytes_per_launch / r["launch_s"] / 1e9
    if args.form == "HIP":
        kname = "stage_f64<8>"
    else:
        kname = f"stage_mp<8, {dgk.tile_shapes('HIP_MP', 7)[dgk.tile_selected('HIP_MP', 7)]}>"
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "HIP" else "f64 state, f32 MFMA contraction+Lift",
        "data": "synthetic (seeded Gaussian blob, seeded constant velocity)",
        "config": {"workload": f"32^3 periodic hexahedra, p=7, {dof} DOF, SSP-RK3 steps, "
                               f"{'plain FP64' if args.form == 'HIP' else 'mixed precision'}",
                   "NeX": wl.NeX, "p": wl.p, "dof": dof, "dt": wl.dt, "form": args.form},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(args.form), "kernel": kname,
                     "avg_launch_us": 1e6 * r["launch_s"], "launches": r["launches"],
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
        "extra": {"q_minmax_after": r["minmax"], "device_timers": r["timers"]},
    }
    if not args.no_extra:
        other = "HIP_MP" if args.form == "HIP" else "HIP"
        _, r2 = time_solver(dgk, inputs, other, args.steps, args.warmup)
        line["extra"][f"c2_{other}"] = {
            "DOF-updates/s": r2["value"], "avg_launch_us": 1e6 * r2["launch_s"],
            "achieved_GBs": BYTES_PER_DOF_UPDATE[other] / 3.0 * dof / r2["launch_s"] / 1e9,
            "ratio_vs_" + args.form: r2["value"] / r["value"]}
        line["extra"]["batched_ensemble"] = batch_leg(dgk, inputs)
    if args.serve:
        from dgk import serving

        sv = serving.run_stream(n=args.serve, rate=args.serve_rate)
        if args.serve_burst:
            cap = serving.run_stream(n=args.serve_burst, rate=1e6, sample=0)
            sv["capacity_burst"] = {k: cap[k] for k in ("requests", "DOF-updates/s",
                                                        "makespan_s", "p50_latency_s",
                                                        "p99_latency_s")}
        line["extra"]["serving"] = sv
        line["extra"]["p50_latency_s"] = sv["p50_latency_s"]
        line["extra"]["p99_latency_s"] = sv["p99_latency_s"]
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps)
        except Exception as exc:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
