"""The C ABI of libdgk.so without a GPU: every declared symbol is exported,
host-side helpers work, and invalid inputs fail with the documented status
codes (CPU).  Solver creation on a host without a HIP device must fail
loudly (DGK_ERR_DEVICE): there is no CPU fallback."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

HEADER = Path(__file__).resolve().parent.parent / "include" / "dgk.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dgk_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(str(dgk.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(dgk._SIGNATURES)


def test_evaltype_names():
    assert dgk.evaltype_id("GENERAL") == dgk.EVAL_GENERAL
    assert dgk.evaltype_id("OPT1    ") == dgk.EVAL_OPT1  # Fortran blank padding
    assert dgk.evaltype_id("HIP") == dgk.EVAL_HIP
    assert dgk.evaltype_id("HIP_MP") == dgk.EVAL_HIP_MP
    with pytest.raises(dgk.DgkError) as ei:
        dgk.evaltype_id("CUDA")
    assert ei.value.code == dgk.ERR_ARG and "Unsupported dgoptr_evaltype" in str(ei.value)


def test_sizes_and_operator_errors():
    assert dgk.element_sizes(7) == dict(Nq=8, Np=512, Nfp=64, NfpTot=384)
    assert dgk.mesh_sizes(32, 32, 32, 7) == dict(Ne=32768, NeA=38912, NhaloNode=393216)
    with pytest.raises(dgk.DgkError) as ei:
        dgk.element_sizes(20)
    assert ei.value.code == dgk.ERR_UNSUPPORTED
    with pytest.raises(dgk.DgkError) as ei:
        dgk.load_operators(3, "/nonexistent")
    assert ei.value.code == dgk.ERR_IO


def _arrays(m, **over):
    a = dict(D1D=m.D, Lift_mat=m.Lift, VMapM=m.VMapM, VMapP=m.VMapP,
             halo_src_map=m.halo_src_map, normal_fn=m.normal_fn, Escale=m.Escale,
             Fscale=m.Fscale)
    a.update(over)
    return a


def test_create_validates_reference_arrays():
    m = oracle.Mesh(2, 2, 2, 2)
    bad = m.VMapM.copy()
    bad[0, 0] += 1
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, **_arrays(m, VMapM=bad))
    assert ei.value.code == dgk.ERR_UNSUPPORTED
    badp = m.VMapP.copy()
    badp[0, 0] = m.Np * m.NeA + 5
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, **_arrays(m, VMapP=badp))
    assert ei.value.code == dgk.ERR_ARG
    esc = m.Escale.copy()
    esc[0, 0, 1] *= 1.5
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, **_arrays(m, Escale=esc))
    assert ei.value.code == dgk.ERR_UNSUPPORTED
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("GENERAL", m.Nq, m.Ne, m.NeA, m.NhaloNode, **_arrays(m))
    assert ei.value.code == dgk.ERR_UNSUPPORTED


@pytest.mark.skipif(dgk.device_count() > 0, reason="host has a HIP device")
def test_no_cpu_fallback_without_device():
    m = oracle.Mesh(2, 2, 2, 2)
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.create("HIP", m.Nq, m.Ne, m.NeA, m.NhaloNode, **_arrays(m))
    assert ei.value.code == dgk.ERR_DEVICE
    with pytest.raises(dgk.DgkError) as ei:
        dgk.Solver.structured("HIP", 4, 4, 4, 3)
    assert ei.value.code == dgk.ERR_DEVICE


def test_input_generator_matches_oracle_stream():
    for seed in (1, 2, 77):
        c = np.array([0, 1, 5, 1 << 20, (1 << 39) + 3], dtype=np.uint64)
        ours = inputs.u01(seed, c)
        ref = [oracle.u01(seed, int(x)) for x in c]
        assert np.array_equal(ours, np.array(ref))


def test_serve_footprint_and_no_device():
    from dgk import serving

    # q (3 buffers; the velocity is one constant per request, held in the
    # geometry record), neighbour table, geometry record and summary
    # partials, rounded to 64 KiB blocks, the same for both classes
    for ne, p, cls in (((4, 4, 4), 2, 0), ((32, 32, 32), 15, 1), ((5, 7, 3), 7, 0)):
        ne_tot = int(np.prod(ne))
        dof = ne_tot * (p + 1) ** 3
        fields = 3 * 8 * dof
        got = serving.footprint(ne, p, cls)
        assert got % 65536 == 0
        assert fields + 24 * ne_tot <= got <= fields + 24 * ne_tot + 8 * 256 + 8224 + 65536 + 256 * 10
    with pytest.raises(dgk.DgkError) as ei:
        serving.footprint((4, 4, 4), 20, 0)
    assert ei.value.code == dgk.ERR_UNSUPPORTED
    if dgk.device_count() == 0:
        with pytest.raises(dgk.DgkError) as ei:
            serving.Server(hbm_budget=1 << 20)
        assert ei.value.code == dgk.ERR_DEVICE


def test_serve_request_stream_ranges():
    reqs = inputs.serve_requests(n=3000, seed=5, rate=4.0)
    assert reqs == inputs.serve_requests(n=3000, seed=5, rate=4.0)
    ne = np.array([r["ne"] for r in reqs])
    assert ne.min() == 4 and ne.max() == 32
    p = np.array([r["p"] for r in reqs])
    assert p.min() == 2 and p.max() == 15 and len(set(p)) == 14
    ns = np.array([r["nstep"] for r in reqs])
    assert ns.min() >= 10 and ns.max() <= 500
    gaps = np.diff([0.0] + [r["arrival"] for r in reqs])
    assert np.all(gaps > 0) and abs(gaps.mean() - 0.25) < 0.02  # Poisson, rate 4/s
    frac = np.mean([r["tol_class"] for r in reqs])
    assert 0.45 < frac < 0.55


def test_tile_shapes_host_side():
    # p <= 7: one, two, four wavefronts of elements, then one wavefront
    # direction by direction; at p <= 4 the run of consecutive elements; last
    # the VALU kernel with FP32 velocity fields
    assert dgk.tile_shapes("HIP_MP", 7) == [1, 2, 4, 1, 1]
    assert dgk.tile_shapes("HIP_MP", 3) == [4, 8, 16, 4, 8, 4]
    assert dgk.kernel_name("HIP_MP", 3, 4) == "stage_run_mp<4, true>"
    assert dgk.kernel_name("HIP_MP", 3, 5) == "stage_f64_mp<4, true, false, true>"
    assert dgk.kernel_name("HIP_MP", 7, 4) == "stage_f64_mp<8, true, false, false>"
    assert dgk.kernel_name("HIP_MP", 7, 3) == "stage_mfma<PrecMixed, 8, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP_MP", 8, 2) == "stage_mfma<PrecMixed, 9, 1, true, false, false, false>"
    assert dgk.tile_shapes("HIP_MP", 8) == [1, 2, 1]
    # p >= 8: one element, two elements, one element direction by direction
    assert dgk.tile_shapes("HIP_MP", 15) == [1, 2, 1]
    # FP64: the two VALU kernels, then the FP64 MFMA kernel's shapes (two
    # elements of p = 15 would need more than 160 KiB of LDS), at p <= 7 last
    # the run kernel
    assert dgk.tile_shapes("HIP", 7) == [1, 1, 1, 2, 4, 1, 1]
    assert dgk.tile_shapes("HIP", 3) == [4, 4, 4, 8, 16, 4, 8]
    assert dgk.kernel_name("HIP", 7, 6) == "stage_run<8, true>"
    assert dgk.kernel_name("HIP", 7, 5) == "stage_mfma<PrecF64, 8, 1, true, false, false, false>"
    assert dgk.tile_shapes("HIP", 15) == [1, 1, 1, 1]
    with pytest.raises(dgk.DgkError) as ei:
        dgk.tile_select("HIP", 15, 4)
    assert ei.value.code == dgk.ERR_ARG
    with pytest.raises(dgk.DgkError) as ei:
        dgk.tile_select("HIP_MP", 15, 3)
    assert ei.value.code == dgk.ERR_ARG
    with pytest.raises(dgk.DgkError) as ei:
        dgk.tile_shapes("GENERAL", 7)
    assert ei.value.code == dgk.ERR_UNSUPPORTED
    dgk.tile_select("HIP_MP", 7, 2)
    assert dgk.tile_selected("HIP_MP", 7) == 2
    dgk.tile_select("HIP_MP", 7, -1)
    assert dgk.tile_selected("HIP_MP", 7) == 0
    # mixed defaults: the VALU kernel (last candidate) at p = 1, two
    # wavefronts at p = 4, four at p = 5, direction by direction at p = 14,
    # one wavefront (or one element) else
    assert [dgk.tile_selected("HIP_MP", p) for p in range(1, 16)] == (
        [5, 0, 0, 1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 2, 0])
    assert dgk.kernel_name("HIP_MP", 1, 5) == "stage_f64_mp<2, true, false, true>"
    dgk.tile_select("HIP_MP", 1, 0)
    assert dgk.tile_selected("HIP_MP", 1) == 0
    dgk.tile_select("HIP_MP", 1, -1)
    assert dgk.tile_selected("HIP_MP", 1) == 5
    # FP64 defaults: the register VALU kernel at p = 4 and 6..13, the
    # direction-by-direction MFMA kernel (the last candidate) at p = 14, 15,
    # the staged VALU kernel else
    assert [dgk.tile_selected("HIP", p) for p in range(1, 16)] == (
        [0, 0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3])
    dgk.tile_select("HIP", 13, 0)
    assert dgk.tile_selected("HIP", 13) == 0
    dgk.tile_select("HIP", 13, -1)
    assert dgk.tile_selected("HIP", 13) == 1


def test_shapes_above_the_operator_files():
    # p = 16..19: one element per workgroup on the MFMA kernels, all at once
    # where its three LDS areas fit (FP64 up to p = 17), then direction by
    # direction
    for p in (16, 17):
        assert dgk.tile_shapes("HIP", p) == [1, 1]
        assert dgk.tile_selected("HIP", p) == 1
    for p in (18, 19):
        assert dgk.tile_shapes("HIP", p) == [1]
        assert dgk.tile_selected("HIP", p) == 0
    for p in (16, 17, 18, 19):
        assert dgk.tile_shapes("HIP_MP", p) == [1, 1]
    # mixed: direction by direction at p = 16, 17, all at once at p = 18, 19
    assert [dgk.tile_selected("HIP_MP", p) for p in (16, 17, 18, 19)] == [1, 1, 0, 0]
    assert dgk.kernel_name("HIP", 16, 0) == "stage_mfma<PrecF64, 17, 1, false, true, false, false>"
    assert dgk.kernel_name("HIP", 16, 1) == "stage_mfma<PrecF64, 17, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP", 17, 1) == "stage_mfma<PrecF64, 18, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP", 19, 0) == "stage_mfma<PrecF64, 20, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP_MP", 19, 0) == (
        "stage_mfma<PrecMixed, 20, 1, false, true, false, false>")
    assert dgk.kernel_name("HIP_MP", 19, 1) == (
        "stage_mfma<PrecMixed, 20, 1, true, false, false, false>")
    assert dgk.kernel_name("HIP", 7, 0) == "stage_f64<8, false, true, false, false>"
    assert dgk.kernel_name("HIP", 7, 1) == "stage_f64<8, true, true, false, false>"
    assert dgk.kernel_name("HIP", 7, 1, axis_aligned=False) == "stage_f64<8, true, false, false, false>"
    assert dgk.kernel_name("HIP", 7, 5) == "stage_mfma<PrecF64, 8, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP", 7, 2) == "stage_mfma<PrecF64, 8, 1, false, true, false, false>"
    assert dgk.kernel_name("HIP_MP", 7, 0) == "stage_mfma<PrecMixed, 8, 1, false, false, false, false>"
    assert dgk.kernel_name("HIP_MP", 11, 0) == "stage_mfma<PrecMixed, 12, 1, false, true, false, false>"
    assert dgk.kernel_name("HIP_MP", 15, 2) == "stage_mfma<PrecMixed, 16, 1, true, false, false, false>"
    assert dgk.kernel_name("HIP", 7, 1, constant_velocity=True) == "stage_f64<8, true, true, true, false>"
    assert dgk.kernel_name("HIP_MP", 7, 0, constant_velocity=True) == (
        "stage_mfma<PrecMixed, 8, 1, false, false, true, false>")
    # several elements per workgroup: the form reading one geometry record
    # through the scalar cache, unless each element has a record of its own
    assert dgk.kernel_name("HIP", 3, 1) == "stage_f64<4, true, true, false, true>"
    assert dgk.kernel_name("HIP", 3, 1, uniform_geometry=False) == (
        "stage_f64<4, true, true, false, false>")
    assert dgk.kernel_name("HIP_MP", 3, 1) == "stage_mfma<PrecMixed, 4, 8, false, true, false, true>"
    assert dgk.element_sizes(19) == dict(Nq=20, Np=8000, Nfp=400, NfpTot=2400)


def test_calibrate_stream_arguments():
    a, r = C.c_double(0.0), C.c_double(0.0)
    # n must be a positive multiple of 2048 (a multiple of both forms' spans per
    # 256-lane workgroup); checked before any device use
    for n in (0, 1000, 2048 + 512):
        assert dgk.lib().dgk_calibrate_stream(n, 1, C.byref(a), C.byref(r)) == dgk.ERR_ARG
    assert dgk.lib().dgk_calibrate_stream(2048, 0, C.byref(a), C.byref(r)) == dgk.ERR_ARG
    assert dgk.lib().dgk_calibrate_stream(2048, 1, None, C.byref(r)) == dgk.ERR_ARG
