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
  extra         the mixed-precision form on the same mesh; the same run with
                the host upload and download inside the clock; the batched
                ensemble (512 seeded instances, Ne/axis in {4,6,8,12,16},
                p in {3,5,7,11}, 50 steps, graph-replayed) in FP64 and mixed
                precision against the plain FP64 path run instance by
                instance; C3 (16^3, p=15, 100 steps) in both forms with the
                mixed form's error against FP64; per-order stage rates of
                both forms (p = 1, 3, ..., 15 at about 1.7e7 DOF) against
                the HBM and MFMA peaks; serving: the first --serve requests of the seeded
                Poisson stream (Ne/axis 4..32, p 2..15, nstep 10..500, half
                of them in the 1e-5 class) at --serve-rate requests/s, with
                p50/p99 latency, DOF-updates/s and a sample of results checked
                against the same solves alone, plus a burst of --serve-burst
                requests for the serving capacity.
"""
from __future__ import annotations

import argparse
import csv
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
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the headline solver on the default tile shape (profiling runs, "
                         "whose serialised launches would distort the autotuner's timings)")
    ap.add_argument("--serve", type=int, default=150,
                    help="requests of the seeded Poisson stream to serve (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=6.0,
                    help="arrival rate of the serving stream, requests/s")
    ap.add_argument("--serve-burst", type=int, default=60,
                    help="requests submitted at once to measure serving capacity (0: skip)")
    ap.add_argument("--per-order-only", type=int, default=0, metavar="P",
                    help="only run --steps steps of the per-order workload at order P "
                         "(periodic box of about 1.7e7 DOF, default tile shape, --form) "
                         "for counter profiles")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (the reference "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--stream-calibration", action="store_true",
                    help="only run the streaming calibration kernel (five arrays read, "
                         "one written) and print its rates")
    ap.add_argument("--batch-only", action="store_true",
                    help="only build the 512-instance ensemble in --form and run --steps "
                         "graph-replayed steps (for kernel traces of a batched step)")
    ap.add_argument("--per-order-table", type=int, nargs="*", default=None, metavar="P",
                    help="only the per-order table, at these orders (all of 1..19 if none given)")
    ap.add_argument("--batch-orders", type=int, nargs="*", default=None, metavar="P",
                    help="with --batch-only: keep only the ensemble's instances of these orders")
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


def copy_rate(nbytes=1 << 30, reps=10):
    """Device-to-device hipMemcpyAsync of `nbytes`, timed with HIP events on the
    null stream: (read + write bytes) / time in GB/s, the practical streaming
    rate of this GPU beside the 8 TB/s datasheet peak."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
        assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0  # warm-up
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def stream_rate(dgk, n=1 << 24, reps=20):
    """dgk_calibrate_stream: a kernel reading five arrays of n doubles and
    writing one (the stage kernels' read:write mix), with 16-byte loads into
    LDS (one per lane and array) and 8-byte loads into registers (one per
    lane and array, the fastest streaming form measured); GB/s of bytes read
    plus written."""
    import ctypes as C

    a, r = C.c_double(0.0), C.c_double(0.0)
    rc = dgk.lib().dgk_calibrate_stream(n, reps, C.byref(a), C.byref(r))
    if rc:
        raise RuntimeError(dgk.lib().dgk_last_error().decode())
    return a.value, r.value


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` on this workload, measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (None when
    no summary covers this kernel)."""
    for f in sorted((ROOT / "profiles").glob("traffic_*_c2_32cube_p7.json")):
        try:
            d = json.loads(f.read_text())
        except (OSError, ValueError):
            continue
        if d.get("kernel") == kernel:
            return d.get("bytes_per_launch")
    return None


def kernel_trace_from_profile(kernel):
    """Mean duration (us) of `kernel` in the rocprofv3 --kernel-trace --stats
    summaries of this workload in profiles/ (None when none covers it)."""
    for f in sorted((ROOT / "profiles").glob("stage_*_c2_32cube_p7_kernel_stats.txt")):
        try:
            lines = f.read_text().splitlines()
        except OSError:
            continue
        for row in csv.reader(l for l in lines if not l.startswith("#")):
            if len(row) > 3 and (kernel + "(") in row[0].replace("dgk::(anonymous namespace)::", ""):
                return float(row[3]) / 1e3
    return None


def time_solver(dgk, inputs, form, steps, warmup, autotune=True):
    wl = inputs.c2_gaussian()
    dof = wl.Ne * wl.Np
    s = dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p)
    s.set_fields(wl.q, wl.u, wl.v, wl.w)
    tune = s.autotune(reps=10) if autotune else {"best": s.tile(), "us_per_launch": None}
    tune["kernel"] = dgk.kernel_name(form, wl.p, tune["best"])
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
                    launches=stage_n, timers=tm, minmax=[qmin, qmax], dof=dof, tune=tune)


def pcie_leg(dgk, inputs, form, steps):
    """The same run with the host boundary inside the clock: q, u, v, w
    uploaded from host memory (dgk_set_field), `steps` steps, q downloaded
    (dgk_get_field).  Reported beside `value`, never as it."""
    wl = inputs.c2_gaussian()
    dof = wl.Ne * wl.Np
    with dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)  # first touch of the device buffers
        s.step(wl.dt, 2)
        s.synchronize()
        t0 = time.perf_counter()
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.synchronize()
        t1 = time.perf_counter()
        s.step(wl.dt, steps)
        s.synchronize()
        t2 = time.perf_counter()
        s.get_field(dgk.FIELD_Q)
        t3 = time.perf_counter()
    return {"form": form, "steps": steps, "DOF-updates/s": dof * steps / (t3 - t0),
            "upload_s": t1 - t0, "steps_s": t2 - t1, "download_s": t3 - t2,
            "host_bytes": 5 * dof * 8}


def batch_leg(dgk, inputs):
    """Batched ensemble (512 instances) in both forms: throughput, every
    instance's mixed-precision error against its FP64 instance, and on the
    p = 7 instances the decomposition of the batched mixed form's rate over
    the plain FP64 path: batching, constant velocity read from the geometry
    record instead of the fields (DGK_CVEL=0 keeps the field reads), and
    precision."""
    ens = inputs.ensemble()
    nstep = ens[0]["nstep"]
    out = {"instances": len(ens), "nstep": nstep}
    dof = np.array([(it["p"] + 1) ** 3 * it["ne"][0] ** 3 for it in ens], dtype=np.float64)
    fields = {}

    def run_batch(form, sub, keep=False):
        b = dgk.Batch(form, sub)
        for k, it in enumerate(sub):
            b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        tiles = b.autotune(reps=10)
        b.step(1)  # graph capture + warm-up
        b.synchronize()
        b.timer()
        t0 = time.perf_counter()
        b.step(nstep)
        b.synchronize()
        t1 = time.perf_counter()
        dev_s, _ = b.timer()
        got = [b.get_field(k) for k in range(len(sub))] if keep else None
        rec = {"DOF-updates/s": float(sum((it["p"] + 1) ** 3 * it["ne"][0] ** 3 for it in sub)
                                      * nstep / (t1 - t0)),
               "seconds": t1 - t0, "device_seconds": dev_s,
               "device_bytes": b.size()["device_bytes"], "tile_shapes": tiles}
        b.close()
        return rec, got

    for form in ("HIP", "HIP_MP"):
        out[form], fields[form] = run_batch(form, ens, keep=True)
        out[form]["bytes_per_dof_update"] = 64
    err = np.array([np.linalg.norm(fields["HIP_MP"][k] - fields["HIP"][k]) /
                    np.linalg.norm(fields["HIP"][k]) for k in range(len(ens))])
    worst = int(err.argmax())
    out["per_instance_mp_vs_fp64"] = {
        "instances": len(ens), "max_rel_l2": float(err.max()), "mean_rel_l2": float(err.mean()),
        "worst_instance": worst, "worst_p": ens[worst]["p"], "worst_ne": ens[worst]["ne"][0],
        "note": "every instance after 51 steps, mixed form against the FP64 form (the FP64 "
                "instances equal their plain-path solves bit for bit, which the parity tests "
                "hold to the oracle)"}
    del fields
    # plain FP64 path, one solver per instance, on the p = 7 instances
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
    solo = p7_dof * nstep / solo_s
    out["plain_fp64_solo_p7"] = {"DOF-updates/s": solo, "seconds": solo_s, "instances": len(p7),
                                 "bytes_per_dof_update": 136}
    sub = [ens[k] for k in p7]
    rates = {}
    for form, cvel in (("HIP", "0"), ("HIP", "1"), ("HIP_MP", "0"), ("HIP_MP", "1")):
        os.environ["DGK_CVEL"] = cvel
        try:
            rec, _ = run_batch(form, sub)
        finally:
            os.environ.pop("DGK_CVEL", None)
        key = f"batched_{form}_p7" + ("" if cvel == "1" else "_velocity_fields")
        rec["bytes_per_dof_update"] = 64 if cvel == "1" else (136 if form == "HIP" else 100)
        rec.pop("tile_shapes", None)
        out[key] = rec
        rates[(form, cvel)] = rec["DOF-updates/s"]
    out["decomposition_p7"] = {
        "batching (plain FP64 solo -> batched FP64, velocity fields)": rates[("HIP", "0")] / solo,
        "constant velocity (batched FP64: fields -> geometry record)": rates[("HIP", "1")] / rates[("HIP", "0")],
        "precision (batched, constant velocity: FP64 -> mixed)": rates[("HIP_MP", "1")] / rates[("HIP", "1")],
        "precision with velocity fields (batched: FP64 -> mixed)": rates[("HIP_MP", "0")] / rates[("HIP", "0")],
        "byte-model cap of the precision gain with velocity fields (136 B -> 100 B)": 136 / 100,
        "batched mixed (constant velocity) / plain FP64 solo": rates[("HIP_MP", "1")] / solo,
    }
    out["speedup_batched_mp_vs_plain_fp64_p7"] = rates[("HIP_MP", "1")] / solo
    return out


def batch_only(dgk, form, steps, orders=None):
    """The 512-instance ensemble (or its instances of the given orders) in one
    form: autotune, capture, then `steps` graph-replayed steps, nothing else
    (the command kernel traces wrap)."""
    from dgk import inputs

    ens = inputs.ensemble()
    if orders:
        ens = [it for it in ens if it["p"] in orders]
    with dgk.Batch(form, ens) as b:
        for k, it in enumerate(ens):
            b.init_gaussian(k, it["centre"], it["sigma"], it["vel"])
        tiles = b.autotune(reps=10)
        b.step(1)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        t1 = time.perf_counter()
    dof = sum((it["p"] + 1) ** 3 * it["ne"][0] ** 3 for it in ens)
    print(json.dumps({"form": form, "steps": steps, "ms_per_step": 1e3 * (t1 - t0) / steps,
                      "DOF-updates/s": dof * steps / (t1 - t0), "tiles": tiles}), flush=True)


def flops_per_dof_update(p):
    """Useful FLOPs of one DOF advanced one SSP-RK3 step: per stage 6 Nq for
    the three contractions, 120/Nq for the six face lifts and fluxes spread
    over the element, 25 for the volume fluxes, the sum and the RK update."""
    nq = p + 1
    return 3.0 * (6 * nq + 25 + 120.0 / nq)


def per_order_box(p):
    """Elements per axis and DOF of the per-order workload at order p."""
    np_ = (p + 1) ** 3
    ne = max(2, round((16_777_216 / np_) ** (1.0 / 3.0)))
    return ne, ne ** 3 * np_


def per_order_only(dgk, form, p, steps):
    """`steps` time steps of the per-order workload at order p, nothing else
    (the command that rocprofv3 counter passes wrap)."""
    ne, dof = per_order_box(p)
    with dgk.Solver.structured(form, ne, ne, ne, p) as s:
        s.init_gaussian((0.3, 0.6, 0.2), 0.1, (0.9, -0.4, 0.3))
        s.step(0.05 * (2.0 / ne) / (p + 1), steps)
        s.synchronize()
    print(json.dumps({"per_order_only": p, "form": form, "ne": ne, "dof": dof, "steps": steps,
                      "kernel": dgk.kernel_name(form, p, dgk.tile_selected(form, p))}), flush=True)


PER_ORDER_ORDERS = tuple(range(1, 20))


def per_order_leg(dgk, orders=PER_ORDER_ORDERS, steps=30, warmup=30):
    """Stage-kernel rate of both forms at every order 1..19 on a periodic box
    of about 1.68e7 DOF (Gaussian blob, constant velocity), default tile
    shapes: achieved algorithmic HBM rate and useful FLOP rate against the
    MI355X peaks (HBM 8 TB/s; MFMA FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s)."""
    out = []
    for p in orders:
        np_ = (p + 1) ** 3
        ne, dof = per_order_box(p)
        for form in ("HIP", "HIP_MP"):
            with dgk.Solver.structured(form, ne, ne, ne, p) as s:
                s.init_gaussian((0.3, 0.6, 0.2), 0.1, (0.9, -0.4, 0.3))
                dt = 0.05 * (2.0 / ne) / np_ ** (1.0 / 3.0)
                # untimed steps first: a launch sequence of a few milliseconds
                # right after set-up runs below the rate the same kernel holds
                # once the device has been busy for a while
                s.step(dt, warmup)
                s.synchronize()
                s.timer_reset()
                s.set_kernel_timing(True)
                s.step(dt, steps)
                s.synchronize()
                sec, n = s.timers()["Stage (fused)"]
            launch = sec / n
            gbs = BYTES_PER_DOF_UPDATE[form] / 3.0 * dof / launch / 1e9
            tfs = flops_per_dof_update(p) / 3.0 * dof / launch / 1e12
            peak = 78.6 if form == "HIP" else 157.3
            shape = dgk.tile_shapes(form, p)[dgk.tile_selected(form, p)]
            out.append({"p": p, "form": form, "ne": ne, "dof": dof, "avg_launch_us": 1e6 * launch,
                        "tile_candidate": dgk.tile_selected(form, p), "elements_per_workgroup": shape,
                        "DOF-updates/s": dof / (3.0 * launch), "alg_GBs": gbs,
                        "frac_hbm": gbs / HBM_PEAK_GBS, "useful_TFLOPs": tfs,
                        "frac_mfma_peak": tfs / peak})
    return out


def c3_leg(dgk, inputs):
    """16^3 elements at p=15 (sin(6 pi x) sin(6 pi y) sin(6 pi z), velocity
    (1,1,1), 100 steps) in both forms: throughput, and the mixed form's error
    against the FP64 device path on the same inputs (the FP64 path is held to
    1e-12 of the reference by the parity tests)."""
    wl = inputs.c3_sine()
    dof = wl.Ne * wl.Np
    out, fields = {"workload": f"16^3 periodic hexahedra, p=15, {dof} DOF, {wl.nstep} steps"}, {}
    for form in ("HIP", "HIP_MP"):
        with dgk.Solver.structured(form, wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
            s.set_fields(wl.q, wl.u, wl.v, wl.w)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(wl.dt, wl.nstep)
            s.synchronize()
            t1 = time.perf_counter()
            fields[form] = s.get_field(dgk.FIELD_Q)
        out[form] = {"DOF-updates/s": dof * wl.nstep / (t1 - t0),
                     "ms_per_step": 1e3 * (t1 - t0) / wl.nstep,
                     "elements_per_workgroup": dgk.tile_shapes(form, wl.p)[dgk.tile_selected(form, wl.p)],
                     "tile_candidate": dgk.tile_selected(form, wl.p)}
    ref, mp = fields["HIP"], fields["HIP_MP"]
    out["mp_rel_l2_vs_fp64"] = float(np.linalg.norm(mp - ref) / np.linalg.norm(ref))
    out["mp_max_rel_vs_fp64"] = float(np.abs(mp - ref).max() / np.abs(ref).max())
    out["speedup_mp_vs_fp64"] = out["HIP_MP"]["DOF-updates/s"] / out["HIP"]["DOF-updates/s"]
    return out


def cpu_baseline(wl, steps):
    """The reference (oracle/_ref, GENERAL) on all host threads: the median of
    three runs of steps/3 SSP-RK3 steps each of the workload.  Two short
    companions: the reference on one core, and the C restatement (oracle.c)
    on all threads, which shows that the restatement is a faithful stand-in
    where the reference is not built.  Threads are left unbound: binding with
    OMP_PLACES=cores serialised the reference's threads in a container whose
    CPU set is a slice of the host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    dof = wl.Ne * wl.Np
    per_run = max(1, steps // 3)
    short = max(1, steps // 10)
    sample = (f"{wl.NeX}^3 elements p={wl.p} ({dof} DOF), median of 3 runs of {per_run} "
              f"SSP-RK3 steps of the same inputs (the GPU line times {wl.nstep}-step runs); "
              f"companions run {short} steps")
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    _, _, port_secs = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, short)
    port = {"value": dof * short / port_secs, "cores": oracle.num_threads(), "steps": short}
    if oracle.reference_available():
        runs = []
        for _ in range(3):
            _, _, secs = oracle.run_reference("time", wl.NeX, wl.NeY, wl.NeZ, wl.p, "GENERAL",
                                              per_run, wl.dt, wl.q, wl.u, wl.v, wl.w,
                                              threads=threads)
            runs.append(dof * per_run / secs)
        _, _, secs1 = oracle.run_reference("time", wl.NeX, wl.NeY, wl.NeZ, wl.p, "GENERAL",
                                           short, wl.dt, wl.q, wl.u, wl.v, wl.w, threads=1)
        return {"value": float(np.median(runs)), "unit": "DOF-updates/s", "cores": threads,
                "kind": "reference", "sample": sample, "runs": runs,
                "one_core": {"value": dof * short / secs1, "steps": short},
                "port_same_inputs": port}
    runs = []
    for _ in range(3):
        _, _, secs = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, per_run)
        runs.append(dof * per_run / secs)
    return {"value": float(np.median(runs)), "unit": "DOF-updates/s",
            "cores": oracle.num_threads(), "kind": "port", "sample": sample, "runs": runs}


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
    if args.stream_calibration:
        a, r = stream_rate(dgk)
        print(json.dumps({"stream_GBs": a, "stream_regs_GBs": r}), flush=True)
        return
    if args.batch_only:
        batch_only(dgk, args.form, args.steps, args.batch_orders)
        return
    if args.per_order_only:
        per_order_only(dgk, args.form, args.per_order_only, args.steps)
        return
    if args.per_order_table is not None:
        for row in per_order_leg(dgk, tuple(args.per_order_table) or PER_ORDER_ORDERS):
            print(json.dumps(row), flush=True)
        return
    from dgk import inputs

    wl, r = time_solver(dgk, inputs, args.form, args.steps, args.warmup,
                        autotune=not args.no_autotune)
    dof = r["dof"]
    bytes_per_launch = BYTES_PER_DOF_UPDATE[args.form] / 3.0 * dof
    achieved = bytes_per_launch / r["launch_s"] / 1e9
    kname = r["tune"]["kernel"]
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
                     "traffic": traffic_from_profile(kname), "kernel": kname,
                     "timing": "one HIP event pair on the solver's stream around the timed "
                               "dgk_step call's back-to-back stage launches",
                     "kernel_trace_avg_us": kernel_trace_from_profile(kname),
                     "avg_launch_us": 1e6 * r["launch_s"], "launches": r["launches"],
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "autotune_us_per_launch": r["tune"]["us_per_launch"]},
        "cpu_baseline": None,
        "extra": {"q_minmax_after": r["minmax"], "device_timers": r["timers"]},
    }
    # the practical streaming rate beside the datasheet peak, and the measured
    # traffic's rate, for reading `frac`
    roof = line["roofline"]
    roof["copy_GBs"] = copy_rate()
    roof["stream_GBs"], roof["stream_regs_GBs"] = stream_rate(dgk)
    if roof["traffic"]:
        roof["traffic_GBs"] = roof["traffic"] / r["launch_s"] / 1e9
    if not args.no_extra:
        other = "HIP_MP" if args.form == "HIP" else "HIP"
        _, r2 = time_solver(dgk, inputs, other, args.steps, args.warmup)
        line["extra"][f"c2_{other}"] = {
            "DOF-updates/s": r2["value"], "avg_launch_us": 1e6 * r2["launch_s"],
            "achieved_GBs": BYTES_PER_DOF_UPDATE[other] / 3.0 * dof / r2["launch_s"] / 1e9,
            "ratio_vs_" + args.form: r2["value"] / r["value"]}
        line["extra"]["pcie_inclusive"] = pcie_leg(dgk, inputs, args.form, args.steps)
        be = batch_leg(dgk, inputs)
        be["decomposition_p7"]["batched mixed (constant velocity) / plain FP64 headline "
                               "(one 32^3 p=7 mesh, velocity fields)"] = (
            be["batched_HIP_MP_p7"]["DOF-updates/s"] / line["value"])
        line["extra"]["batched_ensemble"] = be
        line["extra"]["c3_p15"] = c3_leg(dgk, inputs)
        line["extra"]["per_order"] = per_order_leg(dgk)
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
