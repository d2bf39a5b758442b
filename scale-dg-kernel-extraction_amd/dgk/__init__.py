"""Python binding of libdgk.so (the C ABI declared in include/dgk.h).

The binding is a thin ctypes layer: arrays cross the boundary as the
reference's Fortran arrays (column-major float64 / int32), which map onto
C-contiguous numpy arrays with the axes reversed, e.g. the reference's
q(Np, Ne) is a numpy array of shape (Ne, Np).

There is no CPU fallback: if libdgk.so has not been built, importing the
library raises, and solver creation fails on a host without a HIP device.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libdgk.so"
OPERATOR_DIR = PKG_DIR / "operator_data"

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NOMEM, ERR_IO, ERR_STATE = range(7)
EVAL_GENERAL, EVAL_OPT1, EVAL_HIP, EVAL_HIP_MP = 1, 2, 3, 4
FIELD_Q, FIELD_U, FIELD_V, FIELD_W, FIELD_DQDT = range(5)
TIMER_NAMES = ("Main", "Cal_tend", "Element boundary flux",
               "Volume derivate + surface lift", "Stage (fused)")

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double
_S = C.c_char_p

# name -> (restype, argtypes); mirrors include/dgk.h
_SIGNATURES = {
    "dgk_last_error": (_S, []),
    "dgk_version": (_I, []),
    "dgk_device_count": (_I, [_P]),
    "dgk_evaltype_id": (_I, [_S, _P]),
    "dgk_element_sizes": (_I, [_I, _P, _P, _P, _P]),
    "dgk_load_operators": (_I, [_S, _I, _P, _P, _P]),
    "dgk_make_operators": (_I, [_I, _P, _P, _P]),
    "dgk_write_operators": (_I, [_S, _I]),
    "dgk_mesh_sizes": (_I, [_I, _I, _I, _I, _P, _P, _P]),
    "dgk_mesh_build": (_I, [_I, _I, _I, _I, _P, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P]),
    "dgk_create": (_I, [_I, _I, _L, _L, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dgk_create_structured": (_I, [_I, _I, _I, _I, _I, _S, _D, _D, _D, _P]),
    "dgk_destroy": (_I, [_P]),
    "dgk_solver_info": (_I, [_P, _P, _P, _P, _P]),
    "dgk_device_bytes": (_I, [_P, _P]),
    "dgk_set_field": (_I, [_P, _I, _P, _L]),
    "dgk_get_field": (_I, [_P, _I, _P, _L]),
    "dgk_update_halo": (_I, [_P, _I]),
    "dgk_copy_q0": (_I, [_P]),
    "dgk_cal_elembnd_flux": (_I, [_P, _P]),
    "dgk_cal_dqdt": (_I, [_P, _P]),
    "dgk_cal_tend": (_I, [_P, _P]),
    "dgk_rk_update": (_I, [_P, _I, _D]),
    "dgk_step": (_I, [_P, _D, _I]),
    "dgk_minmax": (_I, [_P, _P, _P]),
    "dgk_minmax_noalloc": (_I, [_P, C.c_int64, _P, _P]),
    "dgk_calibrate_stream": (_I, [_L, _I, _P, _P]),
    "dgk_synchronize": (_I, [_P]),
    "dgk_set_kernel_timing": (_I, [_P, _I]),
    "dgk_timer_report": (_I, [_P, _P, _P]),
    "dgk_timer_reset": (_I, [_P]),
    "dgk_init_gaussian": (_I, [_P, _D, _D, _D, _D, _D, _D, _D]),
    "dgk_batch_create": (_I, [_I, _I, _P, _P, _P, _P, _S, _P]),
    "dgk_batch_destroy": (_I, [_P]),
    "dgk_batch_info": (_I, [_P, _I, _P, _P, _P]),
    "dgk_batch_size": (_I, [_P, _P, _P, _P]),
    "dgk_batch_set_field": (_I, [_P, _I, _I, _P, _L]),
    "dgk_batch_get_field": (_I, [_P, _I, _I, _P, _L]),
    "dgk_batch_init_gaussian": (_I, [_P, _I, _D, _D, _D, _D, _D, _D, _D]),
    "dgk_batch_set_graph": (_I, [_P, _I]),
    "dgk_batch_step": (_I, [_P, _I]),
    "dgk_batch_synchronize": (_I, [_P]),
    "dgk_batch_timer": (_I, [_P, _P, _P]),
    "dgk_tile_shapes": (_I, [_I, _I, _P, _I, _P]),
    "dgk_tile_select": (_I, [_I, _I, _I]),
    "dgk_tile_selected": (_I, [_I, _I, _P]),
    "dgk_autotune": (_I, [_P, _I, _P, _P, _I, _P]),
    "dgk_solver_tile": (_I, [_P, _P]),
    "dgk_batch_autotune": (_I, [_P, _I, _P, _P, _P, _P, _I, _P]),
    "dgk_serve_create": (_I, [_S, _L, _I, _I, _P]),
    "dgk_serve_footprint": (_I, [_P, _I, _I, _P]),
    "dgk_serve_submit": (_I, [_P, _P, _P]),
    "dgk_serve_poll": (_I, [_P, _P, _I, _P, _D]),
    "dgk_serve_set_policy": (_I, [_P, _I, _I]),
    "dgk_serve_get_stats": (_I, [_P, _P]),
    "dgk_serve_destroy": (_I, [_P]),
}

_lib = None


class DgkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dgk error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libdgk.so (built in-tree by `make` in the package directory)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(rc: int) -> None:
    if rc != OK:
        raise DgkError(rc, lib().dgk_last_error().decode())


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags.c_contiguous, "arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().dgk_device_count(C.byref(n)))
    return n.value


def evaltype_id(name: str) -> int:
    i = C.c_int(0)
    _check(lib().dgk_evaltype_id(name.encode(), C.byref(i)))
    return i.value


def tile_shapes(evaltype, p: int) -> list:
    """Elements per workgroup of each tile-shape candidate of a stage kernel."""
    if isinstance(evaltype, str):
        evaltype = evaltype_id(evaltype)
    e, n = np.zeros(8, np.int32), C.c_int(0)
    _check(lib().dgk_tile_shapes(int(evaltype), int(p), _ptr(e), 8, C.byref(n)))
    return e[:n.value].tolist()


def kernel_name(evaltype, p: int, cand: int, axis_aligned: bool = True,
                constant_velocity: bool = False, uniform_geometry: bool = True) -> str:
    """Name of the stage kernel that tile candidate `cand` launches, as a
    kernel trace prints it.  The FP64 VALU kernels come in an axis-aligned and
    a general face-normal form (the structured meshes are axis-aligned); the
    constant-velocity form reads one velocity from the geometry record
    (batches of constant-velocity instances, served requests); the last
    template argument marks a workgroup of several elements reading one
    geometry record for the whole launch (single meshes and served requests,
    not batches, whose instances have records of their own)."""
    if isinstance(evaltype, str):
        evaltype = evaltype_id(evaltype)
    nq = p + 1
    b = lambda x: "true" if x else "false"
    nvalu = 2 if evaltype == EVAL_HIP and nq <= 16 else 0  # FP64 VALU kernels first
    if cand < nvalu:
        gu = uniform_geometry and nq * nq <= 32  # several elements per workgroup
        return (f"stage_f64<{nq}, {b(cand == 1)}, {b(axis_aligned)}, {b(constant_velocity)}, "
                f"{b(gu)}>")
    shapes = tile_shapes(evaltype, p)
    # after the MFMA shapes: the kernel over runs of elements (FP64 at p <= 7,
    # mixed at p <= 4), then the mixed form's VALU kernel (p <= 7)
    has_run = nq <= 8 if evaltype == EVAL_HIP else nq <= 5
    has_valu_mp = evaltype != EVAL_HIP and nq <= 8
    if has_valu_mp and cand == len(shapes) - 1:
        gu = uniform_geometry and nq * nq <= 32
        return f"stage_f64_mp<{nq}, {b(axis_aligned)}, {b(constant_velocity)}, {b(gu)}>"
    if has_run and cand == len(shapes) - 1 - int(has_valu_mp):
        name = "stage_run" if evaltype == EVAL_HIP else "stage_run_mp"
        return f"{name}<{nq}, {b(axis_aligned)}>"
    mfma = shapes[nvalu:len(shapes) - int(has_run) - int(has_valu_mp)]
    mc = cand - nvalu
    epb = mfma[mc]
    # the direction-by-direction shape is the last one; above p = 15 it is the
    # only one where the all-at-once shape's LDS does not fit
    seq = (nq > 16 and len(mfma) == 1) or (
        mc == len(mfma) - 1 and len(mfma) > 1 and (nq * nq <= 64 or epb == 1))
    prec = "PrecF64" if evaltype == EVAL_HIP else "PrecMixed"
    # the direction-by-direction shapes and the mixed one-wavefront shapes
    # test the face normals at run time (instantiated as the general form)
    one_wave = epb * nq * nq <= 64
    ax_runtime = seq or (evaltype != EVAL_HIP and one_wave)
    return (f"stage_mfma<{prec}, {nq}, {epb}, {b(seq)}, {b(axis_aligned and not ax_runtime)}, "
            f"{b(constant_velocity)}, {b(uniform_geometry and epb > 1)}>")


def tile_select(evaltype, p: int, cand: int) -> None:
    if isinstance(evaltype, str):
        evaltype = evaltype_id(evaltype)
    _check(lib().dgk_tile_select(int(evaltype), int(p), int(cand)))


def tile_selected(evaltype, p: int) -> int:
    if isinstance(evaltype, str):
        evaltype = evaltype_id(evaltype)
    c = C.c_int(0)
    _check(lib().dgk_tile_selected(int(evaltype), int(p), C.byref(c)))
    return c.value


def element_sizes(p: int) -> dict:
    v = [C.c_int(0) for _ in range(4)]
    _check(lib().dgk_element_sizes(p, *[C.byref(x) for x in v]))
    return dict(zip(("Nq", "Np", "Nfp", "NfpTot"), (x.value for x in v)))


def load_operators(p: int, directory: str | os.PathLike = OPERATOR_DIR):
    """x1D (Nq,), D1D as numpy (Nq, Nq) with D1D[l, i] = D1D(i, l), and
    Lift_mat as numpy (6, Nq, Nq, Nq) with Lift[f, k, j, i] = Lift_mat(i, j, k, f)."""
    Nq = p + 1
    x = np.zeros(Nq)
    D = np.zeros((Nq, Nq))
    L = np.zeros((6, Nq, Nq, Nq))
    _check(lib().dgk_load_operators(str(directory).encode(), p, _ptr(x), _ptr(D), _ptr(L)))
    return x, D, L


def make_operators(p: int):
    """Closed-form x1D, D1D and Lift_mat of order p, laid out as load_operators."""
    Nq = p + 1
    x = np.zeros(Nq)
    D = np.zeros((Nq, Nq))
    L = np.zeros((6, Nq, Nq, Nq))
    _check(lib().dgk_make_operators(int(p), _ptr(x), _ptr(D), _ptr(L)))
    return x, D, L


def write_operators(directory, p: int) -> None:
    """Write directory/p{p}.dat in the reference's operator format."""
    _check(lib().dgk_write_operators(str(directory).encode(), int(p)))


def mesh_sizes(NeX: int, NeY: int, NeZ: int, p: int) -> dict:
    v = [C.c_int64(0) for _ in range(3)]
    _check(lib().dgk_mesh_sizes(NeX, NeY, NeZ, p, *[C.byref(x) for x in v]))
    return dict(zip(("Ne", "NeA", "NhaloNode"), (x.value for x in v)))


def mesh_build(NeX: int, NeY: int, NeZ: int, p: int, x1D: np.ndarray,
               dom=(1.0, 1.0, 1.0), with_pos: bool = True) -> dict:
    """The reference's mesh arrays (numpy axes reversed w.r.t. Fortran):
    VMapM, VMapP (Ne, NfpTot) int32 1-based; halo_src_map (NhaloNode,);
    normal_fn (3, Ne, NfpTot); Escale (3, Ne, Np); Fscale (Ne, NfpTot);
    pos_en (3, Ne, Np)."""
    es = element_sizes(p)
    ms = mesh_sizes(NeX, NeY, NeZ, p)
    Ne, NfpTot, Np = ms["Ne"], es["NfpTot"], es["Np"]
    out = dict(
        VMapM=np.zeros((Ne, NfpTot), np.int32),
        VMapP=np.zeros((Ne, NfpTot), np.int32),
        halo_src_map=np.zeros(ms["NhaloNode"], np.int32),
        normal_fn=np.zeros((3, Ne, NfpTot)),
        Escale=np.zeros((3, Ne, Np)),
        Fscale=np.zeros((Ne, NfpTot)),
        pos_en=np.zeros((3, Ne, Np)) if with_pos else None,
    )
    _check(lib().dgk_mesh_build(
        NeX, NeY, NeZ, p, _ptr(_f64(x1D)), float(dom[0]), float(dom[1]), float(dom[2]),
        _ptr(out["VMapM"]), _ptr(out["VMapP"]), _ptr(out["halo_src_map"]),
        _ptr(out["normal_fn"]), _ptr(out["Escale"]), _ptr(out["Fscale"]),
        _ptr(out["pos_en"])))
    out.update(ms)
    out.update(es)
    return out


class Solver:
    """A device-resident advection solver (one HIP stream)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        Nq, Ne, NeA, NH = C.c_int(0), C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(lib().dgk_solver_info(handle, C.byref(Nq), C.byref(Ne), C.byref(NeA), C.byref(NH)))
        self.Nq, self.Ne, self.NeA, self.NhaloNode = Nq.value, Ne.value, NeA.value, NH.value
        self.Np = self.Nq ** 3
        self.NfpTot = 6 * self.Nq ** 2

    @classmethod
    def create(cls, evaltype, Nq, Ne, NeA, NhaloNode, D1D, Lift_mat, VMapM, VMapP,
               halo_src_map, normal_fn, Escale, Fscale) -> "Solver":
        """dgk_create from the reference's arrays (advect3d_eq_cal_tend inputs)."""
        if isinstance(evaltype, str):
            evaltype = evaltype_id(evaltype)
        h = C.c_void_p()
        keep = [_f64(D1D), _f64(Lift_mat), _i32(VMapM), _i32(VMapP), _i32(halo_src_map),
                _f64(normal_fn), _f64(Escale), _f64(Fscale)]
        _check(lib().dgk_create(evaltype, Nq, Ne, NeA, NhaloNode, *[_ptr(a) for a in keep],
                                C.byref(h)))
        return cls(h)

    @classmethod
    def structured(cls, evaltype, NeX, NeY, NeZ, p, dom=(1.0, 1.0, 1.0),
                   operator_dir=OPERATOR_DIR) -> "Solver":
        if isinstance(evaltype, str):
            evaltype = evaltype_id(evaltype)
        h = C.c_void_p()
        _check(lib().dgk_create_structured(evaltype, NeX, NeY, NeZ, p,
                                           str(operator_dir).encode(), float(dom[0]),
                                           float(dom[1]), float(dom[2]), C.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            _check(lib().dgk_destroy(self._h))
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

    # fields ---------------------------------------------------------------
    def set_field(self, which: int, f: np.ndarray) -> None:
        a = _f64(f).reshape(-1)
        _check(lib().dgk_set_field(self._h, which, _ptr(a), a.size))

    def get_field(self, which: int, with_halo: bool = False) -> np.ndarray:
        ne = self.NeA if with_halo else self.Ne
        out = np.zeros((ne, self.Np))
        _check(lib().dgk_get_field(self._h, which, _ptr(out), out.size))
        return out

    def set_fields(self, q, u, v, w) -> None:
        for which, f in ((FIELD_Q, q), (FIELD_U, u), (FIELD_V, v), (FIELD_W, w)):
            self.set_field(which, f)

    # reference stage pieces ----------------------------------------------
    def update_halo(self, which: int = FIELD_Q) -> None:
        _check(lib().dgk_update_halo(self._h, which))

    def copy_q0(self) -> None:
        _check(lib().dgk_copy_q0(self._h))

    def cal_elembnd_flux(self, download: bool = True):
        out = np.zeros((self.Ne, self.NfpTot)) if download else None
        _check(lib().dgk_cal_elembnd_flux(self._h, _ptr(out)))
        return out

    def cal_dqdt(self, flux: np.ndarray | None = None) -> None:
        a = None if flux is None else _f64(flux)
        _check(lib().dgk_cal_dqdt(self._h, _ptr(a)))

    def cal_tend(self, download: bool = True):
        out = np.zeros((self.Ne, self.Np)) if download else None
        _check(lib().dgk_cal_tend(self._h, _ptr(out)))
        return out

    def rk_update(self, stage: int, dt: float) -> None:
        _check(lib().dgk_rk_update(self._h, stage, float(dt)))

    def step(self, dt: float, nstep: int) -> None:
        _check(lib().dgk_step(self._h, float(dt), int(nstep)))

    def minmax(self):
        a, b = C.c_double(0), C.c_double(0)
        _check(lib().dgk_minmax(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def synchronize(self) -> None:
        _check(lib().dgk_synchronize(self._h))

    def device_bytes(self) -> int:
        b = C.c_int64(0)
        _check(lib().dgk_device_bytes(self._h, C.byref(b)))
        return b.value

    # timers ----------------------------------------------------------------
    def set_kernel_timing(self, on: bool) -> None:
        _check(lib().dgk_set_kernel_timing(self._h, int(bool(on))))

    def timers(self) -> dict:
        sec = np.zeros(5)
        cnt = np.zeros(5, np.int64)
        _check(lib().dgk_timer_report(self._h, _ptr(sec), _ptr(cnt)))
        return {n: (float(s), int(c)) for n, s, c in zip(TIMER_NAMES, sec, cnt)}

    def timer_reset(self) -> None:
        _check(lib().dgk_timer_reset(self._h))

    def init_gaussian(self, centre, sigma, vel) -> None:
        """Device-side Gaussian blob and constant velocity (structured solvers)."""
        _check(lib().dgk_init_gaussian(self._h, *(float(c) for c in centre), float(sigma),
                                       *(float(c) for c in vel)))

    def tile(self) -> int:
        """Tile-shape candidate this solver's launches use."""
        c = C.c_int(0)
        _check(lib().dgk_solver_tile(self._h, C.byref(c)))
        return c.value

    def autotune(self, reps: int = 10) -> dict:
        """Time every tile shape of this solver's form and order on its mesh and
        keep the fastest for this solver."""
        best, n = C.c_int(0), C.c_int(0)
        us = np.zeros(8)
        _check(lib().dgk_autotune(self._h, int(reps), C.byref(best), _ptr(us), 8, C.byref(n)))
        return dict(best=best.value, us_per_launch=us[:n.value].tolist())


class Batch:
    """Independent periodic boxes advanced together (dgk_batch_*).

    instances: sequence of dicts with keys ne (NeX, NeY, NeZ), p, dt and
    optionally dom (Lx, Ly, Lz)."""

    def __init__(self, evaltype, instances, operator_dir=OPERATOR_DIR):
        if isinstance(evaltype, str):
            evaltype = evaltype_id(evaltype)
        n = len(instances)
        ne = np.array([list(it["ne"]) for it in instances], np.int32).reshape(-1)
        p = np.array([it["p"] for it in instances], np.int32)
        dom = np.array([list(it.get("dom", (1.0, 1.0, 1.0))) for it in instances]).reshape(-1)
        dt = np.array([it["dt"] for it in instances], np.float64)
        h = C.c_void_p()
        _check(lib().dgk_batch_create(evaltype, n, _ptr(ne), _ptr(p), _ptr(dom), _ptr(dt),
                                      str(operator_dir).encode(), C.byref(h)))
        self._h = h
        self.n = n
        self.evaltype = evaltype

    def close(self) -> None:
        if self._h:
            _check(lib().dgk_batch_destroy(self._h))
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

    def info(self, inst: int) -> dict:
        p, ne, dof = C.c_int(0), C.c_int64(0), C.c_int64(0)
        _check(lib().dgk_batch_info(self._h, inst, C.byref(p), C.byref(ne), C.byref(dof)))
        return dict(p=p.value, Ne=ne.value, dof=dof.value)

    def size(self) -> dict:
        n, g, b = C.c_int(0), C.c_int(0), C.c_int64(0)
        _check(lib().dgk_batch_size(self._h, C.byref(n), C.byref(g), C.byref(b)))
        return dict(instances=n.value, groups=g.value, device_bytes=b.value)

    def set_field(self, inst: int, which: int, f) -> None:
        a = _f64(f).reshape(-1)
        _check(lib().dgk_batch_set_field(self._h, inst, which, _ptr(a), a.size))

    def get_field(self, inst: int, which: int = FIELD_Q) -> np.ndarray:
        info = self.info(inst)
        Np = (info["p"] + 1) ** 3
        out = np.zeros((info["Ne"], Np))
        _check(lib().dgk_batch_get_field(self._h, inst, which, _ptr(out), out.size))
        return out

    def init_gaussian(self, inst: int, centre, sigma, vel) -> None:
        _check(lib().dgk_batch_init_gaussian(self._h, inst, *(float(c) for c in centre),
                                             float(sigma), *(float(c) for c in vel)))

    def set_graph(self, on: bool) -> None:
        _check(lib().dgk_batch_set_graph(self._h, int(bool(on))))

    def step(self, nstep: int) -> None:
        _check(lib().dgk_batch_step(self._h, int(nstep)))

    def synchronize(self) -> None:
        _check(lib().dgk_batch_synchronize(self._h))

    def timer(self):
        sec, calls = C.c_double(0), C.c_int64(0)
        _check(lib().dgk_batch_timer(self._h, C.byref(sec), C.byref(calls)))
        return sec.value, calls.value

    def autotune(self, reps: int = 10) -> list:
        """Autotune the tile shape of every order group; returns one record per
        group: order, chosen elements per workgroup, its and the default's
        microseconds per stage launch."""
        m = 16
        p, best, n = np.zeros(m, np.int32), np.zeros(m, np.int32), C.c_int(0)
        ub, ud = np.zeros(m), np.zeros(m)
        _check(lib().dgk_batch_autotune(self._h, int(reps), _ptr(p), _ptr(best), _ptr(ub),
                                        _ptr(ud), m, C.byref(n)))
        form = self.evaltype
        out = []
        for g in range(n.value):
            shapes = tile_shapes(form, int(p[g]))
            out.append(dict(p=int(p[g]), candidate=int(best[g]), epb=shapes[int(best[g])],
                            us_per_launch=float(ub[g]), us_default=float(ud[g])))
        return out
