"""Library rules on the GPU, through the C ABI:
- a recorded run is dropped when what it computed changes (materials, flags,
  sources), so the adjoint cannot mix old and new;
- after an adjoint that recomputed segments, stepping and field reads wait
  for a reset, while the recorded IR stays readable;
- dL/dIR must have the recorded shape;
- a diverging run is reported by virs_synchronize in every precision;
- virs_time_steps times exactly the launches virs_step makes (graphs of any
  length), after an untimed sampling pass;
- checkpoints beyond the device memory fail with an error code, and the
  context stays usable;
- the RMS stop, held off while timed, recorded and adjoint steps run, is in
  effect again afterwards;
- contexts on one GPU share no state."""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, tiny_problem
from oracle import oracle as O
from virs_amd import grid as G
from virs_amd.fdtd import FDTD, FDTDError

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("change", ["materials", "flags", "soft", "hard"])
def test_changes_after_recording_drop_it(gpu, change):
    p = tiny_problem(31, dims=(20, 16, 12), steps=40, sponge=2)
    n = p["steps"]
    _say(f"recording dropped by {change}")
    f = gpu_context(p, "fp64")
    f.forward_record(n, 3)
    dl = np.zeros((len(p["rcv_idx"]), n), np.float32)
    if change == "materials":
        f.set_materials(p["coeffs"][::-1].copy())  # same count
    elif change == "flags":
        f.upload_flags(p["flags"])
    elif change == "soft":
        f.set_soft_sources(p["soft_idx"], p["soft_sig"])
    else:
        f.set_hard_source(None)
    with pytest.raises(FDTDError, match="forward_record first"):
        f.adjoint(dl)
    # a new recording works again
    f.start(n)
    f.forward_record(n, 3)
    assert np.isfinite(f.adjoint(dl)).all()
    f.close()


@pytest.mark.parametrize("ncp", [1, 3])
def test_stepping_after_adjoint(gpu, ncp):
    p = tiny_problem(32, dims=(20, 16, 12), steps=45, sponge=2)
    n = p["steps"]
    _say(f"stepping after an adjoint with {ncp} checkpoints")
    f = gpu_context(p, "fp32")
    f.forward_record(n, ncp)
    ir = f.receivers(True)
    end = f.field(0)
    dl = np.random.default_rng(3).standard_normal(ir.shape).astype(np.float32)
    f.adjoint(dl)
    np.testing.assert_array_equal(f.receivers(True), ir)
    if ncp == 1:
        # nothing was recomputed: the finished run's state is still there
        np.testing.assert_array_equal(f.field(0), end)
    else:
        with pytest.raises(FDTDError, match="virs_start_simulation"):
            f.step(1)
        with pytest.raises(FDTDError, match="virs_start_simulation"):
            f.field(0)
        with pytest.raises(FDTDError, match="virs_start_simulation"):
            f.plane_request(1)
    f.start(n)
    f.step(n)
    np.testing.assert_array_equal(f.receivers(True), ir)
    np.testing.assert_array_equal(f.field(0), end)
    f.close()


def test_dl_dir_shape_is_checked(gpu):
    p = tiny_problem(33, dims=(20, 16, 12), steps=30, sponge=0)
    n, r = p["steps"], len(p["rcv_idx"])
    f = gpu_context(p, "fp32")
    f.forward_record(n, 2)
    for bad in ((r, n - 1), (n, r), (r - 1, n)):
        with pytest.raises(ValueError, match="dl_dir"):
            f.adjoint(np.zeros(bad, np.float32))
    assert np.isfinite(f.adjoint(np.ones((r, n), np.float32))).all()
    f.close()


@pytest.mark.parametrize("case", ["fp32-biquad", "fp64-biquad", "fp16-biquad", "fp32-courant"])
def test_divergence_is_reported(gpu, case):
    """An unstable biquad (poles at 2.32 and 1.68) or a Courant number above
    the 3-D limit (lambda^2 = 0.5) overflows the fields to Inf/NaN; nothing
    faults, and virs_synchronize reports the divergence."""
    prec, kind = case.split("-")
    if kind == "biquad":
        p = tiny_problem(34, dims=(20, 16, 12), steps=1500, sponge=0, hard=False)
        p["coeffs"] = np.tile([1.0, 0.0, 0.0, -4.0, 3.9], (len(p["coeffs"]), 1))
    else:
        p = tiny_problem(35, dims=(20, 16, 12), steps=400, sponge=0, hard=False, lam2=0.5)
    _say(f"divergence {case}, {p['steps']} steps")
    f = gpu_context(p, prec)
    f.step(p["steps"])
    with pytest.raises(FDTDError, match="diverged"):
        f.synchronize()
    # a fresh run clears the flag
    f.start(p["steps"])
    f.synchronize()
    f.close()


def test_stable_run_is_not_flagged(gpu):
    p = tiny_problem(36, dims=(20, 16, 12), steps=300, sponge=2)
    f = gpu_context(p, "fp32")
    f.step(p["steps"])
    f.synchronize()
    f.close()


@pytest.mark.parametrize("n", [20, 47, 65])
def test_timed_steps_are_virs_steps(gpu, n):
    """virs_time_steps: nsample single steps, then n steps as graph replays
    (full 30-step graphs and one graph of the rest); the result equals plain
    single-launch stepping bit for bit, and virs_step with graphs as well."""
    nsample = 8
    p = tiny_problem(37, dims=(24, 18, 14), steps=nsample + n + 5, sponge=3)
    _say(f"timed steps n={n}")
    a = gpu_context(p, "fp32", use_graph=True)
    t = a.time_steps(n, nsample)
    full, tail = divmod(n, 30)
    assert t["sampled_steps"] == nsample
    assert t["graph_launches"] == full + (tail > 0)
    assert t["graph_steps"] == (30 if full else 0)
    assert t["tail_graph_steps"] == tail
    assert t["single_steps"] == 0
    assert t["total_ms"] > 0 and t["dense_ms"] > 0 and t["boundary_ms"] > 0
    b = gpu_context(p, "fp32", use_graph=False)
    b.step(nsample + n)
    c = gpu_context(p, "fp32", use_graph=True)
    c.step(nsample)
    c.step(n)
    assert a.steps_done == b.steps_done == c.steps_done == nsample + n
    for other in (b, c):
        np.testing.assert_array_equal(a.receivers(True), other.receivers(True))
        np.testing.assert_array_equal(a.field(0), other.field(0))
        np.testing.assert_array_equal(a.field(1), other.field(1))
    u = gpu_context(p, "fp32", use_graph=False)
    tu = u.time_steps(n, nsample)
    assert tu["single_steps"] == n and tu["graph_launches"] == 0
    np.testing.assert_array_equal(u.receivers(True), b.receivers(True))
    for f in (a, b, c, u):
        f.close()


def test_checkpoints_beyond_device_memory_fail_with_an_error(gpu):
    """An allocation failure forced from the host: 4000 FP64
    checkpoints of a 256^3 room (2 x 134 MB each, ~1.1 TB) fail at once with
    VIRS_ERR_NOMEM, without touching the device's memory, and the context
    records and differentiates normally afterwards."""
    n = 256
    flags = np.zeros((n, n, n), np.uint8)
    flags[1:-1, 1:-1, 1] = 1
    flags = G.build_shell(flags)
    bidx = G.boundary_list(flags)
    f = FDTD((n, n, n), 0.05, lam2=1 / 3, precision="fp64")
    f.upload_flags(flags)
    f.upload_boundary(bidx)
    f.set_materials([[0.6, -0.2, 0.05, -0.3, 0.1]])
    c = (n // 2 * n + n // 2) * n + n // 2
    f.set_soft_sources([c], np.ones((1, 4000)))
    f.set_receivers([c + 7])
    f.start(4000)
    with pytest.raises(FDTDError, match="checkpoints need"):
        f.forward_record(4000, 4000)
    with pytest.raises(FDTDError, match="forward_record first"):
        f.adjoint(np.zeros((1, 4000), np.float32))
    f.start(40)
    f.forward_record(40, 4)
    g = f.adjoint(np.ones((1, 40), np.float32))
    assert np.isfinite(g).all()
    f.close()


def test_rms_stop_is_restored_after_timing_recording_and_adjoint(gpu):
    """virs_time_steps, virs_forward_record and virs_adjoint run every step
    with the RMS rule (simulator_gpu.cu:303-330) off; a plain run afterwards
    stops at the oracle's step again."""
    p = tiny_problem(2, dims=(28, 20, 16))
    flags = p["flags"]
    nz, ny, nx = flags.shape
    h, steps = 0.05, 3400
    dt = 0.5 * h / (343.0 * np.sqrt(3.0))
    row = G.rbj_highpass(np.float32(1.0 / dt), 440.0, 0.1, 1.0).astype(np.float64)[None]
    ridx = ((nz // 2) * ny + ny // 2 + 2) * nx + nx // 2
    bidx = G.boundary_list(flags)
    ir = O.Problem(flags, h, bidx=bidx, coeffs=row, rcv_idx=[ridx]).run(steps, "f32")["ir"][0]
    stop = O.rms_continue_sequence(ir.astype(np.float32))
    assert 3000 < stop < steps - 100
    f = FDTD((nx, ny, nz), h)
    f.upload_flags(flags)
    f.upload_boundary(bidx)
    f.set_materials(row)
    f.set_receivers([ridx])
    f.set_early_stop(True)

    def stops_at_oracle_step(after):
        _say(f"RMS stop after {after}")
        f.start(steps)
        assert f.step(steps) is False
        assert f.steps_done == stop + 1

    stops_at_oracle_step("nothing")
    f.start(steps)
    t = f.time_steps(steps - 8, 8)  # past the stopping step, all run
    assert t["sampled_steps"] == 8 and f.steps_done == steps
    stops_at_oracle_step("timed steps")
    f.start(steps)
    f.forward_record(steps, 4)
    assert f.steps_done == steps
    stops_at_oracle_step("a recording")
    f.start(steps)
    f.forward_record(steps, 4)
    assert np.isfinite(f.adjoint(np.ones((1, steps), np.float32))).all()
    stops_at_oracle_step("an adjoint")
    f.close()


def test_contexts_interleaved_on_one_gpu(gpu):
    """Three contexts on one GPU (different rooms and storage precisions, with
    and without graphs), stepped in alternation with uneven batches, each
    give the bits of their own run alone: no state is shared between
    contexts.  The reference keeps its constants in process-global
    __constant__ memory uploaded once per process (simulator_gpu.cu:21-36,
    189-212), so a second simulator in one process would reuse the first
    one's grid dimensions."""
    pa = tiny_problem(41, dims=(24, 18, 14), steps=70)
    pb = tiny_problem(42, dims=(20, 16, 13), steps=70, nmat=3)
    pc = tiny_problem(43, dims=(32, 12, 10), steps=70, hard=False)
    _say("three contexts interleaved: fp32 graphs, fp64 direct, fp16 graphs")
    fa = gpu_context(pa, "fp32", use_graph=True)
    fb = gpu_context(pb, "fp64")
    fc = gpu_context(pc, "fp16", use_graph=True)
    for n in (7, 30, 1, 32):
        for f in (fa, fb, fc):
            f.step(n)
    for f in (fa, fb, fc):
        f.synchronize()
    ra = oracle_problem(pa).run(70, "f32", want_fields=True)
    rb = oracle_problem(pb).run(70, "f64", want_fields=True)
    rc = oracle_problem(pc).run16(70, "fp16", stochastic=True, want_fields=True)
    for f, r in ((fa, ra), (fb, rb), (fc, rc)):
        np.testing.assert_array_equal(f.receivers(f64=True), r["ir"])
        np.testing.assert_array_equal(f.field(0), r["p"])
        f.close()
