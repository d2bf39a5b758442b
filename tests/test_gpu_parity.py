"""Parity of the FP64 HIP path ("HIP") with the C oracle, through the C ABI.

The oracle restates the reference GENERAL path and is bit-identical to the
reference Fortran (tests/test_oracle_golden.py).  Tolerance for the FP64
device path: max-norm relative error <= 1e-13 for one tendency evaluation and
<= 1e-12 after a time loop (the FMA-fused device arithmetic differs from the
reference's unfused products by rounding only).
"""
import numpy as np
import pytest

import dgk
import oracle
from dgk import inputs

pytestmark = pytest.mark.gpu

TEND_TOL = 1e-13
LOOP_TOL = 1e-12


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def solver_from_mesh(m, evaltype="HIP", **over):
    arr = dict(D1D=m.D, Lift_mat=m.Lift, VMapM=m.VMapM, VMapP=m.VMapP,
               halo_src_map=m.halo_src_map, normal_fn=m.normal_fn, Escale=m.Escale,
               Fscale=m.Fscale)
    arr.update(over)
    return dgk.Solver.create(evaltype, m.Nq, m.Ne, m.NeA, m.NhaloNode, **arr)


def mesh_for(p):
    return (3, 2, 2) if p <= 6 else (2, 2, 2) if p <= 11 else (2, 1, 2)


@pytest.mark.parametrize("p", range(1, 16))
def test_cal_tend_every_order(p):
    nx, ny, nz = mesh_for(p)
    m = oracle.Mesh(nx, ny, nz, p)
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=100 + p)
    ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        got = s.cal_tend()
    assert maxrel(got, ref) <= TEND_TOL


@pytest.mark.parametrize("p", [1, 3, 7, 8, 15])
def test_elembnd_flux_and_cal_dqdt(p):
    nx, ny, nz = mesh_for(p)
    m = oracle.Mesh(nx, ny, nz, p)
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=200 + p)
    Q, U, V, W = m.field(q), m.field(u), m.field(v), m.field(w)
    flux_ref = m.elembnd_flux(Q, U, V, W)
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        flux = s.cal_elembnd_flux()
        assert maxrel(flux, flux_ref) <= 1e-15
        # cal_dqdt from an externally given flux
        s.cal_dqdt(flux_ref)
        dq = s.get_field(dgk.FIELD_DQDT)
    ref = m.cal_dqdt(Q, U, V, W, flux_ref)
    assert maxrel(dq, ref) <= TEND_TOL


def test_update_halo_matches_reference_order():
    m = oracle.Mesh(3, 2, 4, 4)
    q, u, v, w = inputs.random_fields(3, 2, 4, 4, seed=5)
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        s.update_halo(dgk.FIELD_Q)
        got = s.get_field(dgk.FIELD_Q, with_halo=True)
        gu = s.get_field(dgk.FIELD_U, with_halo=True)
    ref = m.field(q)
    assert np.array_equal(got.reshape(-1)[: m.Np * m.Ne + m.NhaloNode],
                          ref.reshape(-1)[: m.Np * m.Ne + m.NhaloNode])
    assert np.array_equal(gu.reshape(-1)[: m.Np * m.Ne + m.NhaloNode],
                          m.field(u).reshape(-1)[: m.Np * m.Ne + m.NhaloNode])


def test_structured_equals_reference_arrays():
    nx, ny, nz, p = 4, 3, 2, 5
    m = oracle.Mesh(nx, ny, nz, p)
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=9)
    with solver_from_mesh(m) as a, dgk.Solver.structured("HIP", nx, ny, nz, p) as b:
        a.set_fields(q, u, v, w)
        b.set_fields(q, u, v, w)
        assert np.array_equal(a.cal_tend(), b.cal_tend())
        a.step(1e-3, 3)
        b.step(1e-3, 3)
        assert np.array_equal(a.get_field(dgk.FIELD_Q), b.get_field(dgk.FIELD_Q))


def test_node_connectivity_twisted_faces():
    """VMapP that does not pair faces node-for-node (z-boundary faces read the
    transposed node order) runs through the per-node connectivity."""
    nx, ny, nz, p = 2, 2, 3, 3
    m = oracle.Mesh(nx, ny, nz, p)
    Nq, Nfp = m.Nq, m.Nfp
    VP = m.VMapP.copy()
    perm = np.arange(Nfp).reshape(Nq, Nq).T.reshape(-1)
    for e in range(m.Ne):
        for f in (4, 5):
            seg = VP[e, f * Nfp:(f + 1) * Nfp].copy()
            VP[e, f * Nfp:(f + 1) * Nfp] = seg[perm]
    m.VMapP = VP
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=11)
    ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        got = s.cal_tend()
    assert maxrel(got, ref) <= TEND_TOL


def test_per_element_geometry_records():
    """Distinct Escale/Fscale per element (several geometry records)."""
    nx, ny, nz, p = 3, 2, 2, 4
    m = oracle.Mesh(nx, ny, nz, p)
    scale = 1.0 + 0.25 * np.arange(m.Ne)
    m.Escale = np.ascontiguousarray(m.Escale * scale[None, :, None])
    m.Fscale = np.ascontiguousarray(m.Fscale * scale[:, None])
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=12)
    ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        got = s.cal_tend()
    assert maxrel(got, ref) <= TEND_TOL


def test_piecewise_stage_loop_equals_fused_step():
    """copy_q0 / update_halo / cal_tend / rk_update (main.f90:51-75 piece by
    piece) against dgk_step and the oracle loop."""
    nx, ny, nz, p = 3, 3, 2, 3
    m = oracle.Mesh(nx, ny, nz, p)
    q, u, v, w = inputs.random_fields(nx, ny, nz, p, seed=13, vel_scale=0.5)
    dt, nstep = 2e-3, 4
    ref, _, _ = m.run(q, u, v, w, dt, nstep)
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        for _ in range(nstep):
            s.copy_q0()
            for stage in (1, 2, 3):
                s.update_halo(dgk.FIELD_Q)
                s.cal_tend(download=False)
                s.rk_update(stage, dt)
        piece = s.get_field(dgk.FIELD_Q)
    with solver_from_mesh(m) as s:
        s.set_fields(q, u, v, w)
        s.step(dt, nstep)
        fused = s.get_field(dgk.FIELD_Q)
    assert maxrel(piece, ref) <= LOOP_TOL
    assert maxrel(fused, ref) <= LOOP_TOL


def test_c1_full_run_matches_oracle():
    wl = inputs.c1_sine_noise()
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    ref, mm_ref, _ = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, wl.nstep, trace=True)
    with dgk.Solver.structured("HIP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        mm = []
        for _ in range(wl.nstep):
            s.step(wl.dt, 1)
            mm.append(s.minmax())
        got = s.get_field(dgk.FIELD_Q)
    assert maxrel(got, ref) <= LOOP_TOL
    assert np.abs(np.array(mm) - mm_ref).max() <= 1e-12 * np.abs(mm_ref).max()


@pytest.mark.parametrize("nstep", [3])
def test_c2_full_size_short_run(nstep):
    """32^3 elements at p=7 (16.8M DOF): a few steps against the oracle."""
    wl = inputs.c2_gaussian(nstep=nstep)
    m = oracle.Mesh(wl.NeX, wl.NeY, wl.NeZ, wl.p)
    ref, _, _ = m.run(wl.q, wl.u, wl.v, wl.w, wl.dt, nstep)
    with dgk.Solver.structured("HIP", wl.NeX, wl.NeY, wl.NeZ, wl.p) as s:
        s.set_fields(wl.q, wl.u, wl.v, wl.w)
        s.step(wl.dt, nstep)
        got = s.get_field(dgk.FIELD_Q)
    assert maxrel(got, ref) <= LOOP_TOL


@pytest.mark.parametrize("p", range(1, 16))
def test_fp64_tile_shapes_against_oracle(p):
    """Every FP64 tile shape (candidates 0, 1: VALU kernels with v, w staged
    or in registers; then the FP64 MFMA kernel shapes; at p <= 7 last the
    kernel over runs of consecutive elements) holds the tendency and loop
    bars against the oracle; the two VALU kernels and the run kernel give the
    same bits, and so do the MFMA shapes among themselves, also with a partial
    last workgroup (30 elements)."""
    ne = (3, 2, 5) if p <= 7 else (2, 1, 3)
    m = oracle.Mesh(*ne, p)
    q, u, v, w = inputs.random_fields(*ne, p, seed=900 + p, vel_scale=0.8)
    dt = 0.2 * inputs.cfl_dt(inputs.lgl_nodes(p), ne, (0.8, 0.8, 0.8))
    tend_ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    loop_ref, _, _ = m.run(q, u, v, w, dt, 3)
    shapes = dgk.tile_shapes("HIP", p)
    assert len(shapes) >= 3
    loops = []
    try:
        for cand in range(len(shapes)):
            dgk.tile_select("HIP", p, cand)
            with dgk.Solver.structured("HIP", *ne, p) as s:
                s.set_fields(q, u, v, w)
                assert maxrel(s.cal_tend(), tend_ref) <= TEND_TOL, cand
                s.step(dt, 3)
                got = s.get_field(dgk.FIELD_Q)
            assert maxrel(got, loop_ref) <= LOOP_TOL, cand
            loops.append(got)
    finally:
        dgk.tile_select("HIP", p, -1)
    assert np.array_equal(loops[1], loops[0])
    nmfma = len(loops) - 2
    if p <= 7:
        assert dgk.kernel_name("HIP", p, len(loops) - 1).startswith("stage_run<")
        assert np.array_equal(loops[-1], loops[0])
        nmfma -= 1
    for k in range(3, 2 + nmfma):
        assert np.array_equal(loops[k], loops[2]), k


@pytest.mark.parametrize("p", range(1, 8))
def test_run_kernel_uniform_tilted_normals(p):
    """The run kernel's general face-normal form: one geometry record with
    off-axis normals shared by every element (so the launch is not handed to
    another kernel), against the oracle and with the VALU kernel's bits."""
    ne = (4, 3, 2)
    m = oracle.Mesh(*ne, p)
    nf = m.normal_fn.copy()
    Nfp = m.Nfp
    for f in range(6):
        seg = slice(f * Nfp, (f + 1) * Nfp)
        a = int(np.argmax(np.abs(nf[:, 0, f * Nfp])))
        nf[(a + 1) % 3, :, seg] = 0.3
        nf[(a + 2) % 3, :, seg] = -0.2
    m.normal_fn = nf
    q, u, v, w = inputs.random_fields(*ne, p, seed=1500 + p, vel_scale=0.8)
    dt = 0.2 * inputs.cfl_dt(inputs.lgl_nodes(p), ne, (0.8, 0.8, 0.8))
    tend_ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    loop_ref, _, _ = m.run(q, u, v, w, dt, 3)
    run = len(dgk.tile_shapes("HIP", p)) - 1
    got = {}
    try:
        for cand in (0, run):
            dgk.tile_select("HIP", p, cand)
            with solver_from_mesh(m) as s:
                s.set_fields(q, u, v, w)
                tend = s.cal_tend()
                s.step(dt, 3)
                got[cand] = (tend, s.get_field(dgk.FIELD_Q))
    finally:
        dgk.tile_select("HIP", p, -1)
    assert maxrel(got[run][0], tend_ref) <= TEND_TOL
    assert maxrel(got[run][1], loop_ref) <= LOOP_TOL
    assert np.array_equal(got[run][0], got[0][0]) and np.array_equal(got[run][1], got[0][1])


ORDER_MESHES = {
    # plane of 64 elements: the default period 64 gives each XCD one row of 8
    # per plane; zperiod:512 is whole planes, zperiod:128 two rows per turn
    # DGK_ALT=global walks the whole order backwards on every other launch
    # (the default walks each XCD's own range backwards)
    (8, 8, 16): ({}, {"DGK_ORDER": "identity"}, {"DGK_ALT": "0"}, {"DGK_ALT": "global"},
                 {"DGK_ORDER": "identity", "DGK_ALT": "0"}, {"DGK_ORDER": "brick:2,2,4"},
                 {"DGK_ORDER": "zperiod:512"}, {"DGK_ORDER": "zperiod:128"}),
    # plane of 24 elements, 5 planes: eighths of 3 elements; groups of 5
    (4, 6, 5): ({}, {"DGK_ORDER": "identity"}, {"DGK_ORDER": "zperiod:40"}),
    # planes of 15 and 25 elements: uneven eighths (1 or 2, 3 or 4 elements);
    # zperiod:120 deals whole planes of 15
    (3, 5, 8): ({}, {"DGK_ORDER": "identity"}, {"DGK_ALT": "0"}, {"DGK_ALT": "global"},
                {"DGK_ORDER": "zperiod:120"}),
    (5, 5, 3): ({}, {"DGK_ORDER": "identity"}),
    # eighths of 128 elements: the FP64 VALU kernels take the interleaved
    # order too (below 96 elements per eighth they keep the numbering)
    (32, 32, 2): ({}, {"DGK_ORDER": "identity"}, {"DGK_ORDER": "zperiod:2048"}),
}


@pytest.mark.parametrize("ne", list(ORDER_MESHES))
@pytest.mark.parametrize("form", ["HIP", "HIP_MP"])
def test_processing_order_does_not_change_results(form, ne, monkeypatch):
    # plane-numbered periodic meshes, so the default deals an eighth of every
    # plane to each XCD and alternates direction; the element numbering,
    # other periods, a brick order and a single direction must give the same
    # bits, through both the structured and the reference-array constructors
    p, nstep = 3, 6
    x1D = inputs.lgl_nodes(p)
    X, Y, Z = inputs.node_coords(*ne, x1D)
    q = inputs.gaussian_blob(X, Y, Z, (0.4, 0.55, 0.3), 0.12)
    vel = (0.8, -0.45, 0.6)
    u, v, w = (np.full_like(q, c) for c in vel)
    dt = inputs.cfl_dt(x1D, ne, vel)
    m = oracle.Mesh(*ne, p)
    runs = {}
    for env in ORDER_MESHES[ne]:
        monkeypatch.delenv("DGK_ORDER", raising=False)
        monkeypatch.delenv("DGK_ALT", raising=False)
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        for how in ("structured", "arrays"):
            s = (dgk.Solver.structured(form, *ne, p) if how == "structured"
                 else solver_from_mesh(m, form))
            with s:
                s.set_fields(q, u, v, w)
                s.step(dt, nstep)
                runs[(tuple(sorted(env.items())), how)] = s.get_field(dgk.FIELD_Q)
    ref = runs[((), "structured")]
    for key, got in runs.items():
        assert np.array_equal(got, ref), key
    if form == "HIP":
        want, _, _ = m.run(q, u, v, w, dt, nstep)
        assert maxrel(ref, want) <= LOOP_TOL


def tilted_normals(m):
    """normal_fn with off-axis components, constant over each face and
    differing between elements (several geometry records), so the stage
    kernels take their general normal-velocity form
    (VelM = u n1 + v n2 + w n3, mod_advect3d_eq.f90:121-127)."""
    nf = m.normal_fn.copy()
    Nfp = m.Nfp
    for e in range(m.Ne):
        for f in range(6):
            seg = slice(f * Nfp, (f + 1) * Nfp)
            a = int(np.argmax(np.abs(nf[:, e, f * Nfp])))
            nf[(a + 1) % 3, e, seg] = 0.3 + 0.05 * (e % 3)
            nf[(a + 2) % 3, e, seg] = -0.2
    return nf


@pytest.mark.parametrize("p", [1, 3, 5, 7, 8, 12, 15])
def test_general_face_normals_every_tile_shape(p):
    """Face normals off the face axes through the reference-array constructor:
    every FP64 tile shape against the oracle's tendency and time loop, and
    every mixed-precision tile shape to its relative-L2 bar."""
    ne = (3, 2, 2) if p <= 7 else (2, 1, 2)
    m = oracle.Mesh(*ne, p)
    m.normal_fn = tilted_normals(m)
    q, u, v, w = inputs.random_fields(*ne, p, seed=1300 + p, vel_scale=0.8)
    dt = 0.2 * inputs.cfl_dt(inputs.lgl_nodes(p), ne, (0.8, 0.8, 0.8))
    tend_ref = m.cal_tend(m.field(q), m.field(u), m.field(v), m.field(w))
    loop_ref, _, _ = m.run(q, u, v, w, dt, 3)
    try:
        for cand in range(len(dgk.tile_shapes("HIP", p))):
            dgk.tile_select("HIP", p, cand)
            with solver_from_mesh(m) as s:
                s.set_fields(q, u, v, w)
                assert maxrel(s.cal_tend(), tend_ref) <= TEND_TOL, cand
                s.step(dt, 3)
                assert maxrel(s.get_field(dgk.FIELD_Q), loop_ref) <= LOOP_TOL, cand
    finally:
        dgk.tile_select("HIP", p, -1)
    try:
        for cand in range(len(dgk.tile_shapes("HIP_MP", p))):
            dgk.tile_select("HIP_MP", p, cand)
            with solver_from_mesh(m, "HIP_MP") as s:
                s.set_fields(q, u, v, w)
                got = s.cal_tend()
            rel_l2 = float(np.linalg.norm(got - tend_ref) / np.linalg.norm(tend_ref))
            assert rel_l2 <= 1e-5, cand
    finally:
        dgk.tile_select("HIP_MP", p, -1)


@pytest.mark.parametrize("p", [2, 4, 6])
def test_odd_node_count_staging_tail(p):
    """Odd nodes per element (p = 2, 4, 6): the staged fields are copied in
    aligned 16-byte granules, which for an even-numbered element reach 8 bytes
    past it.  A mesh with an odd element count ends on such an element, so the
    last granule of every staged field reads the allocation's tail pad; both
    VALU forms, walking forward and backward over 4 steps, hold the oracle's
    loop bar and give the same bits."""
    ne = (3, 3, 3)
    m = oracle.Mesh(*ne, p)
    q, u, v, w = inputs.random_fields(*ne, p, seed=1300 + p, vel_scale=0.8)
    dt = 0.2 * inputs.cfl_dt(inputs.lgl_nodes(p), ne, (0.8, 0.8, 0.8))
    ref, _, _ = m.run(q, u, v, w, dt, 4)
    got = []
    try:
        for cand in (0, 1):
            dgk.tile_select("HIP", p, cand)
            with dgk.Solver.structured("HIP", *ne, p) as s:
                s.set_fields(q, u, v, w)
                s.step(dt, 4)
                got.append(s.get_field(dgk.FIELD_Q))
            assert maxrel(got[-1], ref) <= LOOP_TOL, cand
    finally:
        dgk.tile_select("HIP", p, -1)
    assert np.array_equal(got[0], got[1])
