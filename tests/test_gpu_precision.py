"""Storage-precision variants.

1. Bit-exact: FP16 and BF16 storage (per-field power-of-two scale, FP32
   arithmetic, nearest-even or stochastic storage rounding) against the
   oracle's restatement of the same storage rules, on small grids.
2. Accuracy against the FP64 run of the same scheme on the 5x4x3 m shoebox
   with 8 receivers (2000 steps); the largest relative L2 error over the
   receivers, measured: FP32 1.9e-5; FP16 4.5e-3 with stochastic rounding,
   5.0e-3 with nearest-even rounding; BF16 3.4e-2 (stochastic).  The FP16
   error is a noise floor of the 11-bit significand: the per-store rounding
   noise accumulates over the room's decay time.  The 1e-3 goal for FP16 is
   not met; the bounds below pin the measured levels.
"""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, rel_l2, tiny_problem

pytestmark = pytest.mark.gpu


def edc_db(ir):
    """Schroeder energy decay curve in dB."""
    e = np.cumsum((ir[::-1].astype(np.float64)) ** 2)[::-1]
    return 10 * np.log10(np.maximum(e / e[0], 1e-300))


def max_edc_dev(ir, ref, floor_db=-60.0):
    a, b = edc_db(ir), edc_db(ref)
    keep = b > floor_db
    return float(np.max(np.abs(a[keep] - b[keep])))


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("seed,hard", [(0, True), (1, False)])
def test_16bit_storage_bit_exact(gpu, fmt, rounding, seed, hard):
    """With a flag-3 cell the stencil selects the hard value; without one
    the zero cells take the update with a zero Laplacian factor."""
    p = tiny_problem(30 + seed, steps=70, hard=hard)
    print(f"16-bit {fmt} {rounding} seed={seed} hard={hard}", flush=True)
    f = gpu_context(p, fmt, rounding=rounding)
    f.step(p["steps"])
    f.synchronize()
    ref = oracle_problem(p).run16(p["steps"], fmt, rounding == "stochastic", want_fields=True)
    np.testing.assert_array_equal(f.receivers(True), ref["ir"])
    np.testing.assert_array_equal(f.field(0), ref["p"])
    np.testing.assert_array_equal(f.field(1), ref["q"])
    assert f.field_exponent(0) == ref["exp"][-1]
    f.close()


def test_fp16_subnormal_neighbours_bit_exact(gpu):
    """Early in a run the wavefront's leading cells hold values far below the
    field's maximum, so some stored FP16 values are subnormal.  The stencil
    adds its y-neighbours to the sum straight from the stored pairs; the
    step after such a field still equals the oracle's bit for bit.  The test
    first checks that the oracle's field at that point does hold subnormal
    stored values next to updated cells."""
    p = tiny_problem(31, steps=70, hard=False)
    n = 10
    ref = oracle_problem(p).run16(n, "fp16", True, want_fields=True)
    raw = ref["p"] * 2.0 ** int(ref["exp"][-1])
    sub = (raw != 0) & (np.abs(raw) < 2.0 ** -14)
    upd = p["flags"] != 1
    next_to = np.zeros_like(sub)
    next_to[:, 1:, :] |= sub[:, :-1, :] & upd[:, 1:, :]
    next_to[:, :-1, :] |= sub[:, 1:, :] & upd[:, :-1, :]
    assert next_to.sum() >= 5, "no subnormal stored value is a y-neighbour of an updated cell"
    after = oracle_problem(p).run16(n + 1, "fp16", True, want_fields=True)
    f = gpu_context(p, "fp16")
    f.step(n)
    np.testing.assert_array_equal(f.field(0), ref["p"])
    f.step(1)
    f.synchronize()
    np.testing.assert_array_equal(f.field(0), after["p"])
    np.testing.assert_array_equal(f.field(1), after["q"])
    np.testing.assert_array_equal(f.receivers(True), after["ir"])
    f.close()


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
@pytest.mark.parametrize("nx", [23, 26, 28, 40])
def test_16bit_odd_widths(gpu, fmt, nx):
    """Row widths that force 1-, 2-, 4- and 8-wide vectors in 16-bit
    storage, so the half-byte cell codes are read half a byte, one, two or
    four bytes at a time (with Nx = 23 and Ny = 15 every other plane starts
    on an odd cell), bit-exact with the oracle in both roundings."""
    p = tiny_problem(12, dims=(nx, 15, 13), steps=40)
    for rounding in ("stochastic", "nearest"):
        print(f"16-bit odd width {fmt} nx={nx} {rounding}", flush=True)
        f = gpu_context(p, fmt, rounding=rounding)
        f.step(p["steps"])
        f.synchronize()
        ref = oracle_problem(p).run16(p["steps"], fmt, rounding == "stochastic", want_fields=True)
        np.testing.assert_array_equal(f.receivers(True), ref["ir"])
        np.testing.assert_array_equal(f.field(0), ref["p"])
        np.testing.assert_array_equal(f.field(1), ref["q"])
        f.close()


@pytest.fixture(scope="module")
def shoebox_runs(gpu):
    from virs_amd import rooms
    from virs_amd.fdtd import setup
    prep = rooms.prepare(rooms.shoebox_5x4x3(receivers=8))
    out = {}
    for prec, rnd in (("fp64", "stochastic"), ("fp32", "stochastic"), ("fp16", "stochastic"),
                      ("bf16", "stochastic"), ("fp16", "nearest")):
        print(f"shoebox 8 receivers {prec} {rnd}", flush=True)
        f = setup(prep, prec, rounding=rnd)
        f.step(prep.scenario.steps)
        f.synchronize()
        out[prec if rnd == "stochastic" else prec + "_rne"] = f.receivers(True)
        f.close()
    return out


def test_fp32_vs_fp64(shoebox_runs):
    ref = shoebox_runs["fp64"]
    errs = [rel_l2(shoebox_runs["fp32"][r], ref[r]) for r in range(len(ref))]
    print("fp32 rel L2 per receiver", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-4


# The goal for FP16 storage is <= 1e-3 relative IR error against FP64.  It is
# not met: each 16-bit store rounds with a relative error of ~2^-11/sqrt(12),
# which the undamped leapfrog accumulates like a random walk, and the measured
# levels on this 2000-step shoebox are 4.5e-3 (FP16 stochastic), 5.0e-3 (FP16
# nearest) and 3.4e-2 (BF16).  The bounds below sit just above those
# measured levels, so a regression in the 16-bit path fails them; they do not
# restate the goal.
@pytest.mark.parametrize("key,bound", [("fp16", 6e-3), ("fp16_rne", 1e-2), ("bf16", 5e-2)])
def test_16bit_error_vs_fp64(shoebox_runs, key, bound):
    ref = shoebox_runs["fp64"]
    run = shoebox_runs[key]
    errs = [rel_l2(run[r], ref[r]) for r in range(len(ref))]
    devs = [max_edc_dev(run[r], ref[r]) for r in range(len(ref))]
    print(f"{key} rel L2 per receiver", ["%.2e" % e for e in errs],
          "max EDC dev dB", ["%.3f" % d for d in devs], flush=True)
    assert np.all(np.isfinite(run))
    assert max(errs) <= bound


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
@pytest.mark.parametrize("hard", [True, False])
def test_fused_and_list_boundary_agree_in_scaled_storage(gpu, prec, hard):
    """Fusing the filter into the stencil rounds the corrected boundary value
    once instead of twice; the two paths stay within a few storage ulps,
    with and without a hard-source cell."""
    p = tiny_problem(9, steps=80, hard=hard)
    a = gpu_context(p, prec, fuse=True)
    b = gpu_context(p, prec, fuse=False)
    a.step(p["steps"])
    b.step(p["steps"])
    e = rel_l2(a.receivers(True), b.receivers(True))
    assert e <= (2e-3 if prec == "fp16" else 3e-2)
    a.close()
    b.close()


def test_write_and_read_field_round_trip(gpu):
    """16-bit storage keeps wall and border cells at zero (its stencil relies
    on it); every other cell round-trips to the format's precision."""
    p = tiny_problem(10, steps=1)
    rng = np.random.default_rng(0)
    nx, ny, nz = p["dims"]
    vals = rng.standard_normal((nz, ny, nx)) * 1e-3
    zero = p["flags"] == 1
    zero[0, :, :] = zero[-1, :, :] = zero[:, 0, :] = zero[:, -1, :] = zero[:, :, 0] = zero[:, :, -1] = True
    for prec, tol in (("fp64", 0), ("fp32", 1e-7), ("fp16", 1e-3), ("bf16", 8e-3)):
        f = gpu_context(p, prec)
        f.write_field(0, vals)
        got = f.field(0)
        want = np.where(zero, 0.0, vals) if prec in ("fp16", "bf16") else vals
        assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(vals))
        if prec in ("fp16", "bf16"):
            assert np.all(got[zero] == 0.0)
        f.close()


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
@pytest.mark.parametrize("k", [-40, 30])
def test_16bit_inputs_scaled_by_a_power_of_two_shift_only_the_exponents(gpu, fmt, k):
    """Every step of the 16-bit scheme commutes with scaling by 2^k: the
    stored bits stay the same and each field's exponent moves by -k, so the
    IR and the fields scale exactly.  Both sources are scaled: the soft
    signals and the hard cell's pulse (10 for n < 10, given as a signal).
    The scaled run is also bit-exact against the oracle's run of the scaled
    room, at exponents near the +-60 clamp's far ends (about 50 for 2^-40)."""
    p = tiny_problem(46, steps=60, hard=True)
    n = p["steps"]
    pulse = np.where(np.arange(n) < 10, 10.0, 0.0)
    q = dict(p)
    q["soft_sig"] = p["soft_sig"] * 2.0 ** k
    print(f"16-bit {fmt} sources scaled by 2^{k}", flush=True)
    runs = []
    for prob, hs in ((p, pulse), (q, pulse * 2.0 ** k)):
        f = gpu_context(prob, fmt)
        f.set_hard_source(hs)
        f.start(n)
        f.step(n)
        f.synchronize()
        runs.append((f.receivers(True), f.field(0), f.field_exponent(0)))
        f.close()
    (ir0, p0, e0), (ir1, p1, e1) = runs
    np.testing.assert_array_equal(ir1, ir0 * 2.0 ** k)
    np.testing.assert_array_equal(p1, p0 * 2.0 ** k)
    assert e1 == e0 - k
    pb = oracle_problem(q)
    pb.hard_sig = np.ascontiguousarray(pulse * 2.0 ** k)
    ref = pb.run16(n, fmt, True, want_fields=True)
    np.testing.assert_array_equal(ir1, ref["ir"])
    np.testing.assert_array_equal(p1, ref["p"])
    assert ref["exp"].max() <= 60 and ref["exp"].min() >= -60
