"""Seeded random rooms against the oracle, bit for bit: widths that give
every vector width and chunk alignment, all four precisions, both storage
roundings, sponge on and off, with and without a hard cell, 1-3 materials,
fused and unfused FP32/FP64, and several stencil grid sizes."""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, tiny_problem

pytestmark = pytest.mark.gpu


def _case(i):
    rng = np.random.default_rng(9000 + i)
    nx = int(rng.choice([17, 20, 23, 26, 30, 33, 40, 48, 64, 72, 96, 130, 136, 200, 264, 520]))
    prec = ["fp32", "fp64", "fp16", "bf16"][i % 4]
    return dict(dims=(nx, int(rng.integers(10, 17)), int(rng.integers(10, 15))), prec=prec,
                rounding=str(rng.choice(["stochastic", "nearest"])), sponge=int(rng.choice([0, 3])),
                hard=bool(rng.integers(0, 2)), nmat=int(rng.integers(1, 4)),
                fuse=bool(rng.integers(0, 2)) and prec in ("fp32", "fp64"),
                dense_blocks=int(rng.choice([0, 1, 5, 37])), steps=int(rng.integers(20, 45)))


@pytest.mark.parametrize("i", range(16))
def test_random_room(gpu, i):
    c = _case(i)
    print(f"random room {i}: {c}", flush=True)
    p = tiny_problem(700 + i, dims=c["dims"], nmat=c["nmat"], sponge=c["sponge"], hard=c["hard"],
                     steps=c["steps"])
    kw = dict(fuse=c["fuse"], dense_blocks=c["dense_blocks"])
    if c["prec"] in ("fp16", "bf16"):
        kw["rounding"] = c["rounding"]
    f = gpu_context(p, c["prec"], **kw)
    f.step(c["steps"])
    ir, fld = f.receivers(True), f.field(0)
    f.close()
    pb = oracle_problem(p)
    if c["prec"] in ("fp32", "fp64"):
        ref = pb.run(c["steps"], "f32" if c["prec"] == "fp32" else "f64", want_fields=True)
    else:
        ref = pb.run16(c["steps"], c["prec"], c["rounding"] == "stochastic", want_fields=True)
    np.testing.assert_array_equal(ir, ref["ir"])
    np.testing.assert_array_equal(fld, ref["p"])


def _case16(i):
    rng = np.random.default_rng(51000 + i)
    nx = int(rng.choice([18, 24, 40, 64, 96, 130, 136, 264, 520, 1032]))
    return dict(dims=(nx, int(rng.integers(10, 16)), int(rng.integers(10, 14))), prec=["fp16", "bf16"][i % 2],
                rounding=["stochastic", "nearest"][(i // 2) % 2], sponge=int(rng.choice([0, 3])),
                hard=bool(rng.integers(0, 2)), nmat=int(rng.integers(1, 4)), steps=int(rng.integers(25, 60)),
                dense_blocks=int(rng.choice([0, 0, 3, 37])))


@pytest.mark.parametrize("i", range(24))
def test_random_16bit_room(gpu, i):
    """16-bit storage across both formats and roundings, sponge and hard
    cells on and off, and widths up to two staged halo vectors per thread:
    the update's folded output scale and FP16's mixed q term against the
    oracle's physical-unit arithmetic, bit for bit."""
    c = _case16(i)
    print(f"random 16-bit room {i}: {c}", flush=True)
    p = tiny_problem(9100 + i, dims=c["dims"], nmat=c["nmat"], sponge=c["sponge"], hard=c["hard"],
                     steps=c["steps"])
    f = gpu_context(p, c["prec"], rounding=c["rounding"], dense_blocks=c["dense_blocks"])
    f.step(c["steps"])
    ir, fld = f.receivers(True), f.field(0)
    f.close()
    ref = oracle_problem(p).run16(c["steps"], c["prec"], c["rounding"] == "stochastic", want_fields=True)
    np.testing.assert_array_equal(ir, ref["ir"])
    np.testing.assert_array_equal(fld, ref["p"])
