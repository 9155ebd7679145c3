"""The dispatch adapter (include/virs_dispatch.hpp): the reference's 16
dispatch functions (simulator_dispatch.h:4-25) on the C ABI, driven in
Simulator's call order through a stand-in VectorSpace
(tests/adapter/dispatch_driver.cpp), checked against the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import tiny_problem
from oracle import oracle as O
from virs_amd import grid as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "adapter", "_build", "libdispatch_driver.so")

# simulator_dispatch.h:4-25
REFERENCE_DISPATCH = [
    "fdtd_setup", "fdtd_step", "fdtd_cleanup", "fdtd_start_simulation", "initPressureSphere",
    "buildSpongeLayer", "updateAllGridFromGPU", "updateCurrentGridFromGPU", "updateGPUFromGrid",
    "uploadNormalsToGPU", "uploadPZetaToGPU", "uploadBoundaryIndicesToGPU", "allocFilterStates",
    "allocFilterCoeffs", "uploadFilterCoeffsToGPU",
]


def driver():
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is missing: run __graft_entry__.build() first")
    L = C.CDLL(DRIVER)
    f = L.virs_drive_reference_sequence
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_void_p, C.c_int64,
                  C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def run_sequence(flags, h, coeffs5, steps, receiver, sphere=None, pressure=0.0):
    nz, ny, nx = flags.shape
    os.environ["VIRS_RECEIVER"] = ",".join(str(int(v)) for v in receiver)
    flags = np.ascontiguousarray(flags, np.uint8)
    co = np.ascontiguousarray(coeffs5, np.float32)
    out = dict(flags=np.zeros_like(flags), ir=np.zeros(max(steps, 1), np.float32),
               p=np.zeros(flags.shape, np.float32), q=np.zeros(flags.shape, np.float32),
               r=np.zeros(flags.shape, np.float32))
    sp = None if sphere is None else np.ascontiguousarray(sphere, np.int64)
    run = C.c_int64(0)
    pct = C.c_float(0)
    nb = C.c_int64(0)
    rc = driver()(flags.ctypes.data, nx, ny, nz, h, co.ctypes.data, steps,
                  None if sp is None else sp.ctypes.data, pressure, out["flags"].ctypes.data,
                  out["ir"].ctypes.data, out["p"].ctypes.data, out["q"].ctypes.data,
                  out["r"].ctypes.data, C.byref(run), C.byref(pct), C.byref(nb))
    assert rc == 0, "adapter sequence failed or leaked a host buffer / context"
    out["steps_run"], out["percent"], out["nb"] = run.value, pct.value, nb.value
    out["ir"] = out["ir"][: run.value]
    return out


def test_macro_defines_every_reference_dispatch_function():
    text = open(os.path.join(ROOT, "include", "virs_dispatch.hpp")).read()
    macro = text[text.index("#define VIRS_DEFINE_DISPATCH"):]
    defined = re.findall(r"(?:void|bool)\s+(\w+)\s*\(SPACE", macro)
    assert sorted(defined) == sorted(REFERENCE_DISPATCH)


def test_driver_library_loads():
    assert driver() is not None


def _room(seed):
    """Pre-shell voxel flags (0 air, 1 wall, 3 hard source) of a small room."""
    p = tiny_problem(seed, dims=(28, 20, 16))
    f = p["flags"].copy()
    f[f == 2] = 0
    return f, p


def _reference_row(h):
    # Simulator's material row (simulator.cpp:955-964): RBJ high-pass 440 Hz, Q 0.1
    dt = 0.5 * h / (343.0 * np.sqrt(3.0))
    return G.rbj_highpass(np.float32(1.0 / dt), 440.0, 0.1, 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_simulator_sequence_matches_oracle(gpu, seed):
    flags, p = _room(seed)
    nz, ny, nx = flags.shape
    h = 0.05
    steps = 90
    row = _reference_row(h)
    rcv = (nx // 3, ny // 2, nz // 2)
    print(f"adapter sequence seed={seed} dims={flags.shape}", flush=True)
    out = run_sequence(flags, h, row, steps, rcv)
    shell = G.build_shell(flags)
    np.testing.assert_array_equal(out["flags"], shell)
    bidx = G.boundary_list(shell)
    assert out["nb"] == len(bidx)
    assert out["steps_run"] == steps
    ridx = (rcv[2] * ny + rcv[1]) * nx + rcv[0]
    pb = O.Problem(shell, h, bidx=bidx, coeffs=np.asarray(row, np.float32).astype(np.float64)[None],
                   rcv_idx=[ridx])
    ref = pb.run(steps, "f32", want_fields=True)
    np.testing.assert_array_equal(out["ir"], ref["ir"][0].astype(np.float32))
    np.testing.assert_array_equal(out["p"], ref["p"].astype(np.float32))
    np.testing.assert_array_equal(out["q"], ref["q"].astype(np.float32))
    # p_next after the reference's pointer rotation holds p^{n-2}
    ref2 = pb.run(steps - 1, "f32", want_fields=True)
    np.testing.assert_array_equal(out["r"], ref2["q"].astype(np.float32))


@pytest.mark.gpu
def test_rms_rule_stops_where_the_oracle_stops(gpu):
    flags, p = _room(2)
    nz, ny, nx = flags.shape
    h = 0.05
    steps = 3400
    row = _reference_row(h)
    rcv = (nx // 2, ny // 2 + 2, nz // 2)
    print("adapter RMS stop", flush=True)
    out = run_sequence(flags, h, row, steps, rcv)
    shell = G.build_shell(flags)
    ridx = (rcv[2] * ny + rcv[1]) * nx + rcv[0]
    pb = O.Problem(shell, h, bidx=G.boundary_list(shell),
                   coeffs=np.asarray(row, np.float32).astype(np.float64)[None], rcv_idx=[ridx])
    ir = pb.run(steps, "f32")["ir"][0].astype(np.float32)
    stop = O.rms_continue_sequence(ir)
    assert stop > 3000, "the room must ring down inside the run for this test to bite"
    assert out["steps_run"] == stop + 1
    np.testing.assert_array_equal(out["ir"], ir[: stop + 1])
    assert 0.0 < out["percent"]


@pytest.mark.gpu
def test_pressure_sphere_initial_condition(gpu):
    flags, p = _room(3)
    nz, ny, nx = flags.shape
    h = 0.05
    row = _reference_row(h)
    sphere = (nx // 2, ny // 2, nz // 2, 3)
    out = run_sequence(flags, h, row, 0, (1, 1, 1), sphere=sphere, pressure=0.75)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inside = (x - sphere[0]) ** 2 + (y - sphere[1]) ** 2 + (z - sphere[2]) ** 2 <= sphere[3] ** 2
    np.testing.assert_array_equal(out["p"], np.where(inside, np.float32(0.75), np.float32(0)))
    # 40 steps from the sphere: the oracle started from the same p^0
    steps = 40
    out = run_sequence(flags, h, row, steps, (nx // 3, ny // 2, nz // 2), sphere=sphere, pressure=0.75)
    shell = G.build_shell(flags)
    ridx = (nz // 2 * ny + ny // 2) * nx + nx // 3
    pb = O.Problem(shell, h, bidx=G.boundary_list(shell),
                   coeffs=np.asarray(row, np.float32).astype(np.float64)[None], rcv_idx=[ridx],
                   init_p=np.where(inside, 0.75, 0.0).reshape(-1))
    ref = pb.run(steps, "f32", want_fields=True)
    np.testing.assert_array_equal(out["ir"], ref["ir"][0].astype(np.float32))
    np.testing.assert_array_equal(out["p"], ref["p"].astype(np.float32))


def _baffled_hall():
    """Pre-shell flags of a 128x96x96 room (1.18 M cells): a walled hall
    with x-normal baffles every 6 cells whose gaps alternate between the two
    y ends (a maze that absorbs fast through its boundary cells), and one
    hard-source voxel at (10, 48, 47)."""
    nx, ny, nz = 128, 96, 96
    f = np.zeros((nz, ny, nx), np.uint8)
    f[2, 2:ny - 2, 2:nx - 2] = 1
    f[nz - 3, 2:ny - 2, 2:nx - 2] = 1
    f[2:nz - 2, 2, 2:nx - 2] = 1
    f[2:nz - 2, ny - 3, 2:nx - 2] = 1
    f[2:nz - 2, 2:ny - 2, 2] = 1
    f[2:nz - 2, 2:ny - 2, nx - 3] = 1
    for k, x in enumerate(range(16, nx - 8, 6)):
        f[2:nz - 2, 2:ny - 2, x] = 1
        if k % 2:
            f[2:nz - 2, 2:12, x] = 0
        else:
            f[2:nz - 2, ny - 12:ny - 2, x] = 0
    f[47, 48, 10] = 3
    return f


@pytest.mark.gpu
def test_reference_default_scheme_with_rms_stop(gpu):
    """The reference's own scheme end to end through the dispatch adapter:
    its cell size h = 343/88200 and dt (lambda^2 = 1/12, simulator_gpu.cu:196),
    its material row (RBJ high-pass 440 Hz, Q 0.1 at fs = 1/dt,
    simulator.cpp:942-967; a real pole at 0.99909), the flag-3 hard pulse
    (simulator_gpu.cu:214-218), one receiver and the live RMS stop
    (simulator_gpu.cu:303-330), on a 1.18 M-cell room where the rule fires
    well after step 3000 (the one-pole recursion runs for ~900 steps).
    The GPU IR equals the FP64-state oracle bit for bit and is within 1e-4
    relative L2 of the reference-literal oracle (float filter state, the
    reference's arithmetic), and all three stop at the same step."""
    flags = _baffled_hall()
    nz, ny, nx = flags.shape
    h = G.reference_cell_size()
    assert abs(h - 343.0 / 88200.0) < 1e-9
    row = _reference_row(h)
    rcv = (4, 63, 70)
    steps = 4500
    print(f"reference default scheme, {nx}x{ny}x{nz}, {steps} steps requested", flush=True)
    out = run_sequence(flags, h, row, steps, rcv)
    shell = G.build_shell(flags)
    np.testing.assert_array_equal(out["flags"], shell)
    bidx = G.boundary_list(shell)
    ridx = (rcv[2] * ny + rcv[1]) * nx + rcv[0]
    co = np.asarray(row, np.float32).astype(np.float64)[None]
    O.set_threads(os.cpu_count() or 1)
    n_or = 4000
    ir = O.Problem(shell, h, bidx=bidx, coeffs=co, rcv_idx=[ridx]).run(n_or, "f32")["ir"][0]
    lit = O.Problem(shell, h, bidx=bidx, coeffs=co, rcv_idx=[ridx], state_f32=True).run(n_or, "f32")["ir"][0]
    stop = O.rms_continue_sequence(ir.astype(np.float32))
    stop_lit = O.rms_continue_sequence(lit.astype(np.float32))
    print(f"oracle stop {stop}, reference-literal stop {stop_lit}, GPU ran {out['steps_run']}", flush=True)
    assert 3300 < stop < n_or - 1, "the rule must fire well after step 3000 inside the oracle run"
    assert stop_lit == stop
    assert out["steps_run"] == stop + 1
    np.testing.assert_array_equal(out["ir"], ir[: stop + 1].astype(np.float32))
    rel = np.linalg.norm(out["ir"] - lit[: stop + 1]) / np.linalg.norm(lit[: stop + 1])
    print(f"GPU vs reference-literal oracle: rel L2 {rel:.2e} over {stop + 1} steps", flush=True)
    assert rel <= 1e-4
    assert 0.0 < out["percent"]


@pytest.mark.gpu
def test_shipped_room_at_reference_defaults(gpu):
    """The reference's default room (assets/room_with_source2.obj as
    tests/golden/room_with_source2.npz holds it) at its defaults through the
    dispatch adapter: h = 343/88200 on the 1337 x 900 x 1079 grid (1.30 G
    cells, 15.1 M list entries), the RBJ row, the flag-3 pulse of the source
    mesh and the fixed receiver (334, 225, 270), 2000 steps (the pulse
    reaches the receiver after ~1600).  The shell the adapter builds has the
    recorded digest, the IR is finite and reaches the receiver, and two runs
    are bit-identical.  A 128^3 crop around the source then runs bit-exact
    against the oracle."""
    import hashlib
    z = np.load(os.path.join(ROOT, "tests", "golden", "room_with_source2.npz"))
    mesh = G.mesh_from_arrays(z["tri"], z["mat"])
    h = G.reference_cell_size()
    dims = tuple(int(v) for v in z["dims"])
    pre = G.voxelize(mesh, z["origin"], h, dims)[0]
    row = _reference_row(h)
    steps = 2000
    runs = []
    for k in range(2):
        print(f"shipped room run {k}: {dims}, {steps} steps", flush=True)
        out = run_sequence(pre, h, row, steps, (334, 225, 270))
        assert out["steps_run"] == steps and out["nb"] == int(z["nb"])
        if k == 0:
            assert hashlib.sha256(out["flags"].tobytes()).hexdigest() == str(z["flags_sha256"])
        runs.append((out["ir"].copy(), out["p"]))
        del out
    ir, p = runs[0]
    assert np.isfinite(ir).all() and np.abs(ir).max() > 0
    np.testing.assert_array_equal(runs[1][0], ir)
    np.testing.assert_array_equal(runs[1][1], p)
    del runs, p
    # two 128^3 crops bit-exact against the oracle: around the source mesh
    # (the flag-3 pulse, no walls nearby) and around the receiver (walls and
    # boundary filters, driven by a soft Gaussian-derivative source)
    from virs_amd import rooms
    from virs_amd.fdtd import FDTD
    shell = G.build_shell(pre)
    src = np.argwhere(pre == 3)
    row64 = np.asarray(row, np.float32).astype(np.float64)[None]
    O.set_threads(os.cpu_count() or 1)
    n_crop = 300
    for name, lo in (("source", (src.min(0) + src.max(0)) // 2 - 64), ("receiver", np.array([270, 225, 334]) - 64)):
        crop = np.ascontiguousarray(shell[lo[0]:lo[0] + 128, lo[1]:lo[1] + 128, lo[2]:lo[2] + 128])
        bidx = G.boundary_list(crop)
        air = np.flatnonzero(crop.reshape(-1) == 0)
        rcv = air[[len(air) // 5, len(air) // 2, (4 * len(air)) // 5]]
        soft = [] if name == "source" else [int(air[(2 * len(air)) // 5])]
        sig = rooms.gauss_deriv(n_crop)[None] if soft else None
        if name == "source":
            assert (crop == 3).any()
        else:
            assert (crop == 2).any() and not (crop == 3).any()
        print(f"{name} crop at {lo.tolist()}: nb {len(bidx)}, {n_crop} steps", flush=True)
        f = FDTD((128, 128, 128), h)
        f.upload_flags(crop)
        f.upload_boundary(bidx)
        f.set_materials(row64)
        if soft:
            f.set_soft_sources(soft, sig)
        f.set_receivers(rcv)
        f.start(n_crop)
        f.step(n_crop)
        gir, gp = f.receivers(True), f.field(0)
        f.close()
        ref = O.Problem(crop, h, bidx=bidx, coeffs=row64, rcv_idx=rcv, soft_idx=soft,
                        soft_sig=sig).run(n_crop, "f32", want_fields=True)
        assert np.abs(ref["p"]).max() > 0
        np.testing.assert_array_equal(gir, ref["ir"])
        np.testing.assert_array_equal(gp, ref["p"])
