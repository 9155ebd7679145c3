"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: FP32 and FP64 bit-exact against the oracle restatement of the same
scheme (both sides evaluate the same expressions without FP contraction);
FP32 against the reference-literal oracle (FP32 filter state, the
reference's own arithmetic) within 1e-4 relative L2 over the whole IR.
"""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, rel_l2, tiny_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_tiny_bit_exact(gpu, seed, precision, fuse):
    p = tiny_problem(seed)
    _say(f"tiny seed={seed} {precision} fuse={fuse} dims={p['dims']} nb={len(p['bidx'])}")
    f = gpu_context(p, precision, fuse=fuse)
    f.step(p["steps"])
    f.synchronize()
    ir = f.receivers(f64=True)
    fld = f.field(0)
    prev = f.field(1)
    ref = oracle_problem(p).run(p["steps"], "f32" if precision == "fp32" else "f64", want_fields=True)
    np.testing.assert_array_equal(ir, ref["ir"])
    np.testing.assert_array_equal(fld, ref["p"])
    np.testing.assert_array_equal(prev, ref["q"])
    f.close()


def test_constants_match_oracle(gpu):
    p = tiny_problem(0)
    f = gpu_context(p)
    dt, c2dt2, ih2, k = f.constants
    odt, oc, oi, ok = O.reference_constants(p["h"], dt=p["dt"])
    assert (dt, c2dt2, ih2, k) == (odt, oc, oi, ok)
    # default dt is the reference's 0.5f*h/(c*sqrtf(3.f))
    from virs_amd.fdtd import FDTD
    g = FDTD((8, 8, 8), 343.0 / 88200.0)
    assert g.constants == O.reference_constants(343.0 / 88200.0)
    g.close()
    f.close()


def test_graph_replay_matches_direct(gpu):
    p = tiny_problem(3, steps=95)
    a = gpu_context(p, use_graph=False)
    b = gpu_context(p, use_graph=True)
    a.step(p["steps"])
    # uneven batches: graph replays of 30 steps plus single launches
    b.step(31)
    b.step(64)
    a.synchronize()
    b.synchronize()
    np.testing.assert_array_equal(a.receivers(True), b.receivers(True))
    np.testing.assert_array_equal(a.field(0), b.field(0))
    a.close()
    b.close()


def test_repeatable_and_launch_shape_invariant(gpu):
    """Jacobi-style update: repeated runs and different stencil grid sizes
    (one block marching everything, odd splits, one plane per block) give
    bitwise identical results (the race check)."""
    p = tiny_problem(4, dims=(32, 20, 40), steps=80)
    outs = []
    for nblk in (0, 0, 1, 7, 40, 100000):
        f = gpu_context(p, dense_blocks=nblk)
        f.step(p["steps"])
        outs.append((f.receivers(True), f.field(0)))
        f.close()
    for ir, fld in outs[1:]:
        np.testing.assert_array_equal(ir, outs[0][0])
        np.testing.assert_array_equal(fld, outs[0][1])


@pytest.mark.parametrize("nx", [23, 26, 30, 64, 136])
def test_odd_widths(gpu, nx):
    """Row widths that force 1-, 2- and 4-wide vectors along x, and planes
    that span several 256-thread chunks."""
    p = tiny_problem(5, dims=(nx, 15, 13), steps=40)
    for prec in ("fp32", "fp64"):
        f = gpu_context(p, prec)
        f.step(p["steps"])
        ref = oracle_problem(p).run(p["steps"], "f32" if prec == "fp32" else "f64", want_fields=True)
        np.testing.assert_array_equal(f.receivers(True), ref["ir"])
        np.testing.assert_array_equal(f.field(0), ref["p"])
        f.close()


def test_reference_hard_source_pulse(gpu):
    """Flag-3 cells hold 10 for steps 0-9 then 0 (simulator_gpu.cu:214-218),
    and a receiver on one reads exactly that."""
    p = tiny_problem(6, steps=20)
    src = int(np.flatnonzero(p["flags"].reshape(-1) == 3)[0])
    p["rcv_idx"] = np.array([src], np.int64)
    f = gpu_context(p)
    f.set_hard_source(None)
    f.step(20)
    ir = f.receivers(True)[0]
    np.testing.assert_array_equal(ir, np.where(np.arange(20) < 10, 10.0, 0.0))
    f.close()


def test_hard_source_signal(gpu):
    p = tiny_problem(7, steps=30)
    sig = np.sin(np.arange(30) * 0.3)
    f = gpu_context(p)
    f.set_hard_source(sig)
    f.step(30)
    pb = oracle_problem(p)
    pb.hard_sig = sig
    ref = pb.run(30, "f32", want_fields=True)
    np.testing.assert_array_equal(f.receivers(True), ref["ir"])
    np.testing.assert_array_equal(f.field(0), ref["p"])
    f.close()


def test_shoebox_2000_steps_vs_reference_scheme(gpu):
    """The 5x4x3 m shoebox (160x132x108, 2000 steps): FP32 IR within 1e-4
    relative L2 of the reference-literal oracle (FP32 filter state), and
    bit-exact against the FP64-state FP32 oracle."""
    from virs_amd import rooms
    from virs_amd.fdtd import setup
    prep = rooms.prepare(rooms.shoebox_5x4x3())
    _say(f"shoebox nb={len(prep.room.bidx)}")
    f = setup(prep, "fp32")
    f.step(prep.scenario.steps)
    ir = f.receivers(True)
    f.close()
    r = prep.room
    O.set_threads(16)
    kw = dict(bidx=r.bidx, bmat=r.bmat, coeffs=prep.scenario.coeffs, soft_idx=prep.src_idx,
              soft_sig=prep.src_sig, rcv_idx=prep.rcv_idx, sponge_width=16, sponge_sigma=0.5)
    lit = O.Problem(r.flags, prep.scenario.h, dt=prep.dt, state_f32=True, **kw).run(2000, "f32")["ir"]
    same = O.Problem(r.flags, prep.scenario.h, dt=prep.dt, **kw).run(2000, "f32")["ir"]
    err = rel_l2(ir, lit)
    _say(f"shoebox fp32 vs reference-literal rel L2 = {err:.3e}")
    assert err <= 1e-4
    np.testing.assert_array_equal(ir, same)


@pytest.mark.parametrize("use_graph", [True, False])
def test_device_rms_stop_in_one_batch(gpu, use_graph):
    """The RMS rule (simulator_gpu.cu:303-330) runs on the device: a single
    3400-step call, replayed as graphs or launched step by step, stops at the
    oracle's step, and the steps after it are skipped on the device (the
    final field is the one right after the stopping step)."""
    from virs_amd import grid as G
    from virs_amd.fdtd import FDTD
    p = tiny_problem(2, dims=(28, 20, 16))
    flags = p["flags"]
    nz, ny, nx = flags.shape
    h, steps = 0.05, 3400
    dt = 0.5 * h / (343.0 * np.sqrt(3.0))
    row = G.rbj_highpass(np.float32(1.0 / dt), 440.0, 0.1, 1.0).astype(np.float64)[None]
    ridx = ((nz // 2) * ny + ny // 2 + 2) * nx + nx // 2
    bidx = G.boundary_list(flags)
    pb = O.Problem(flags, h, bidx=bidx, coeffs=row, rcv_idx=[ridx])
    ir = pb.run(steps, "f32")["ir"][0].astype(np.float32)
    stop = O.rms_continue_sequence(ir)
    assert 3000 < stop < steps - 100
    _say(f"device RMS stop graph={use_graph}, oracle stop {stop}")
    f = FDTD((nx, ny, nz), h, use_graph=use_graph)
    f.upload_flags(flags)
    f.upload_boundary(bidx)
    f.set_materials(row)
    f.set_receivers([ridx])
    f.set_early_stop(True)
    f.start(steps)
    assert f.step(steps) is False
    assert f.steps_done == stop + 1
    np.testing.assert_array_equal(f.receivers()[0], ir[: stop + 1])
    after = pb.run(stop + 1, "f32", want_fields=True)
    np.testing.assert_array_equal(f.field(0), after["p"].astype(np.float32))
    assert f.percent_to_target_rms > 1.0
    f.close()


@pytest.mark.parametrize("prec,nx", [("fp32", 768), ("fp32", 2048), ("fp16", 1536), ("bf16", 2560)])
def test_chunks_ending_mid_row(gpu, prec, nx):
    """Stencil chunks of 256 vectors that end mid-row (Nx = 768 FP32,
    1536 16-bit) or are shorter than a row (Nx = 2048 FP32, 2560 16-bit),
    so waves cross row ends and x-neighbours cross chunks: the oracle's
    results bit for bit, for more than one grid size."""
    p = tiny_problem(15, dims=(nx, 9, 10), steps=30)
    _say(f"chunks ending mid-row {prec} nx={nx}")
    outs = []
    for nblk in (0, 7):
        f = gpu_context(p, prec, dense_blocks=nblk)
        f.step(p["steps"])
        outs.append((f.receivers(True), f.field(0)))
        f.close()
    pb = oracle_problem(p)
    if prec == "fp32":
        ref = pb.run(p["steps"], "f32", want_fields=True)
    else:
        ref = pb.run16(p["steps"], prec, True, want_fields=True)
    np.testing.assert_array_equal(outs[0][0], ref["ir"])
    np.testing.assert_array_equal(outs[0][1], ref["p"])
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    np.testing.assert_array_equal(outs[1][1], outs[0][1])


@pytest.mark.parametrize("prec,nx", [("fp32", 384), ("fp32", 512), ("fp32", 516), ("fp32", 1024),
                                     ("fp32", 1028), ("fp16", 768), ("fp16", 1032), ("fp16", 2048),
                                     ("fp16", 2056), ("bf16", 1536), ("fp64", 256), ("fp64", 258),
                                     ("fp64", 514)])
def test_rows_staged_through_lds_and_not(gpu, prec, nx):
    """The stencil stages each plane through LDS with one halo vector per
    thread when a row has at most 128 vectors (FP32 rows up to 512 cells,
    16-bit up to 1024, FP64 up to 256), with two up to 256 vectors, and
    loads the rows +-Nx itself above that.  Each side of both limits, with
    chunks that end mid-row (Nx = 384 FP32, 768 FP16, 1536 BF16) and with
    several chunks per block (a 7-block grid), gives the oracle's results
    bit for bit."""
    p = tiny_problem(16, dims=(nx, 9, 10), steps=30)
    _say(f"rows staged or not {prec} nx={nx}")
    outs = []
    for nblk in (0, 7):
        f = gpu_context(p, prec, dense_blocks=nblk)
        f.step(p["steps"])
        outs.append((f.receivers(True), f.field(0)))
        f.close()
    pb = oracle_problem(p)
    if prec in ("fp32", "fp64"):
        ref = pb.run(p["steps"], "f32" if prec == "fp32" else "f64", want_fields=True)
    else:
        ref = pb.run16(p["steps"], prec, True, want_fields=True)
    for ir, fld in outs:
        np.testing.assert_array_equal(ir, ref["ir"])
        np.testing.assert_array_equal(fld, ref["p"])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fields_written_mid_run_keep_the_filter_state(gpu, precision):
    """Fields written between steps (virs_write_field_f64, the reference's
    updateGPUFromGrid, simulator_dispatch.cpp:70-81) replace p^n and p^{n-1}
    and leave the boundary filter's state as it was, as the reference's
    device filter_states are left.  Writing the fields back unchanged gives
    the bits of an uninterrupted run; writing them halved gives the bits of
    the oracle continued from the halved fields with its filter state
    carried over (state_io)."""
    p = tiny_problem(44, steps=70)
    n1, n = 35, p["steps"]
    prec = "f32" if precision == "fp32" else "f64"
    _say(f"mid-run field write {precision}")
    straight = gpu_context(p, precision)
    straight.step(n)
    straight.synchronize()
    same = gpu_context(p, precision, use_graph=True)
    same.step(n1)
    P, Q = same.field(0), same.field(1)
    same.write_field(0, P)
    same.write_field(1, Q)
    same.step(n - n1)
    same.synchronize()
    np.testing.assert_array_equal(same.receivers(f64=True), straight.receivers(f64=True))
    np.testing.assert_array_equal(same.field(0), straight.field(0))
    half = gpu_context(p, precision)
    half.step(n1)
    half.write_field(0, 0.5 * half.field(0))
    half.write_field(1, 0.5 * half.field(1))
    half.step(n - n1)
    half.synchronize()
    o = oracle_problem(p)
    o.state_io = np.zeros(4 * max(1, len(p["bidx"])), np.float64)
    first = o.run(n1, prec, want_fields=True)
    o.init_p = np.ascontiguousarray(0.5 * first["p"].reshape(-1))
    o.init_q = np.ascontiguousarray(0.5 * first["q"].reshape(-1))
    second = o.run(n - n1, prec, n0=n1, want_fields=True)
    got = half.receivers(f64=True)
    np.testing.assert_array_equal(got[:, :n1], first["ir"])
    np.testing.assert_array_equal(got[:, n1:], second["ir"])
    np.testing.assert_array_equal(half.field(0), second["p"])
    for f in (straight, same, half):
        f.close()
