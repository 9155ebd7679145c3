"""Storage-precision variants against the FP64 run of the same scheme.

The shoebox workload with 8 receivers (the 5x4x3 m room, 2000 steps): FP16
storage with per-field power-of-two scales and FP32 arithmetic must stay
within 1e-3 relative L2 of the FP64 GPU run at every receiver; BF16 is
measured and bounded loosely (its 8-bit mantissa cannot meet 1e-3).  The
FP64 GPU run itself is bit-exact with the FP64 oracle (test_gpu_parity).
"""
import numpy as np
import pytest

from helpers import gpu_context, rel_l2, tiny_problem

pytestmark = pytest.mark.gpu


def edc_db(ir):
    """Schroeder energy decay curve in dB."""
    e = np.cumsum((ir[::-1].astype(np.float64)) ** 2)[::-1]
    return 10 * np.log10(np.maximum(e / e[0], 1e-300))


def max_edc_dev(ir, ref, floor_db=-60.0):
    a, b = edc_db(ir), edc_db(ref)
    keep = b > floor_db
    return float(np.max(np.abs(a[keep] - b[keep])))


@pytest.fixture(scope="module")
def shoebox_runs(gpu):
    from virs_amd import rooms
    from virs_amd.fdtd import setup
    prep = rooms.prepare(rooms.shoebox_5x4x3(receivers=8))
    out = {}
    for prec in ("fp64", "fp32", "fp16", "bf16"):
        print(f"shoebox 8 receivers {prec}", flush=True)
        f = setup(prep, prec)
        f.step(prep.scenario.steps)
        f.synchronize()
        out[prec] = f.receivers(True)
        f.close()
    return out


def test_fp32_vs_fp64(shoebox_runs):
    ref = shoebox_runs["fp64"]
    errs = [rel_l2(shoebox_runs["fp32"][r], ref[r]) for r in range(len(ref))]
    print("fp32 rel L2 per receiver", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-4


def test_fp16_within_1e3_of_fp64(shoebox_runs):
    ref = shoebox_runs["fp64"]
    errs = [rel_l2(shoebox_runs["fp16"][r], ref[r]) for r in range(len(ref))]
    devs = [max_edc_dev(shoebox_runs["fp16"][r], ref[r]) for r in range(len(ref))]
    print("fp16 rel L2 per receiver", ["%.2e" % e for e in errs],
          "max EDC dev dB", ["%.3f" % d for d in devs], flush=True)
    assert max(errs) <= 1e-3


def test_bf16_bounded(shoebox_runs):
    ref = shoebox_runs["fp64"]
    errs = [rel_l2(shoebox_runs["bf16"][r], ref[r]) for r in range(len(ref))]
    devs = [max_edc_dev(shoebox_runs["bf16"][r], ref[r]) for r in range(len(ref))]
    print("bf16 rel L2 per receiver", ["%.2e" % e for e in errs],
          "max EDC dev dB", ["%.3f" % d for d in devs], flush=True)
    assert np.all(np.isfinite(shoebox_runs["bf16"]))
    assert max(errs) <= 5e-2


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_scaled_storage_tracks_fp64_on_tiny_grid(gpu, prec):
    p = tiny_problem(8, steps=120)
    ref = gpu_context(p, "fp64")
    ref.step(p["steps"])
    f = gpu_context(p, prec)
    f.step(p["steps"])
    e = rel_l2(f.receivers(True), ref.receivers(True))
    print(f"tiny {prec} rel L2 {e:.2e}, exponent {f.field_exponent(0)}", flush=True)
    assert e <= (2e-3 if prec == "fp16" else 3e-2)
    # the stored field's scale keeps its maximum inside the format's range
    fld = f.field(0)
    assert np.max(np.abs(fld)) * 2.0 ** f.field_exponent(0) < 65504
    f.close()
    ref.close()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_fused_and_list_boundary_agree_in_scaled_storage(gpu, prec):
    """Fusing the filter into the stencil rounds the corrected boundary value
    once instead of twice; the two paths stay within a few storage ulps."""
    p = tiny_problem(9, steps=80)
    a = gpu_context(p, prec, fuse=True)
    b = gpu_context(p, prec, fuse=False)
    a.step(p["steps"])
    b.step(p["steps"])
    e = rel_l2(a.receivers(True), b.receivers(True))
    assert e <= (2e-3 if prec == "fp16" else 3e-2)
    a.close()
    b.close()


def test_write_and_read_field_round_trip(gpu):
    p = tiny_problem(10, steps=1)
    rng = np.random.default_rng(0)
    nx, ny, nz = p["dims"]
    vals = rng.standard_normal((nz, ny, nx)) * 1e-3
    for prec, tol in (("fp64", 0), ("fp32", 1e-7), ("fp16", 1e-3), ("bf16", 8e-3)):
        f = gpu_context(p, prec)
        f.write_field(0, vals)
        got = f.field(0)
        assert np.max(np.abs(got - vals)) <= tol * np.max(np.abs(vals))
        f.close()
