#!/usr/bin/env python3
"""Benchmark of the DG advection hot path on one MI355X.

Workload (default): the plain FP64 path ("HIP") on the 32^3-element periodic
cube at p=7 (16,777,216 DOF), velocity U[-1,1]^3 and a Gaussian blob from seed
2, dt from CFL 0.1.  A "step" is one full SSP-RK3 time step of the whole mesh
(3 fused stage kernels).  Inputs are resident in HBM before the timed region.

Metric: DOF-updates/s = Ne*Np*K / t for K timed steps.

Printed (one JSON line, rank 0): the driver contract fields plus
  roofline     -- the fused stage kernel's algorithmic HBM bytes per launch
                  (136/3 B per DOF: q and u,v,w read, q written, q0 read on
                  stages 2 and 3) over its average launch duration, from HIP
                  events recorded on the solver's stream around every launch;
  cpu_baseline -- the reference Fortran GENERAL path (oracle/_ref, OpenMP on
                  the host cores) timed on a bounded sample of the same mesh;
  extra        -- per-kernel timings and the other forms when run.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--form", default="HIP", choices=["HIP", "HIP_MP"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=6,
                    help="time steps of the bounded CPU baseline sample")
    return ap.parse_args()


def traffic_from_profile():
    """HBM bytes per launch measured by rocprofv3 PMC runs (profiles/)."""
    f = ROOT / "profiles" / "traffic_stage_f64_c2.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(wl, steps):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    dof = wl.Ne * wl.Np
    sample = f"{wl.NeX}^3 elements p={wl.p}, {steps} SSP-RK3 steps, same inputs"
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
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "multi-GPU sharding is a design (see DESIGN.md), not a run mode"}))
        sys.exit(2)
    import dgk
    from dgk import inputs

    wl = inputs.c2_gaussian()
    dof = wl.Ne * wl.Np
    s = dgk.Solver.structured(args.form, wl.NeX, wl.NeY, wl.NeZ, wl.p)
    s.set_fields(wl.q, wl.u, wl.v, wl.w)
    s.step(wl.dt, args.warmup)
    s.synchronize()

    # timed region: K steps, HIP events around every stage launch
    s.timer_reset()
    s.set_kernel_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(wl.dt, args.steps)
    s.synchronize()
    t1 = time.perf_counter()
    s.set_kernel_timing(False)
    tm = s.timers()
    secs = t1 - t0
    value = dof * args.steps / secs
    stage_s, stage_n = tm["Stage (fused)"]
    avg_launch = stage_s / max(stage_n, 1)
    bytes_per_launch = BYTES_PER_DOF_UPDATE[args.form] / 3.0 * dof
    achieved = bytes_per_launch / avg_launch / 1e9
    qmin, qmax = s.minmax()
    s.close()

    line = {
        "metric": "DOF-updates/s plain FP64 vs batched mixed-prec; p50/p99 solve latency",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "HIP" else "f64 state / f32 MFMA contractions",
        "data": "synthetic (seeded Gaussian blob, seeded constant velocity)",
        "config": {"workload": f"32^3 periodic hexahedra, p=7, {dof} DOF, SSP-RK3, "
                               f"form {args.form}",
                   "NeX": wl.NeX, "p": wl.p, "dof": dof, "dt": wl.dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(),
                     "kernel": f"stage_{'f64' if args.form == 'HIP' else 'mp'}<8>",
                     "avg_launch_us": 1e6 * avg_launch, "launches": stage_n,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
        "extra": {"q_minmax_after": [qmin, qmax], "device_timers": tm},
    }
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps)
        except Exception as exc:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
