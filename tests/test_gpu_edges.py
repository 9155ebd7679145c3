"""Edge cases of the step through the C ABI, against the oracle: a room with
no boundary list at all, and the point tables at their limits (64 soft
sources and 64 receivers, some of them on list cells, two receivers on one
cell)."""
import numpy as np
import pytest

from helpers import oracle_problem, tiny_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run_oracle(pb, steps, precision, rounding="stochastic"):
    if precision in ("fp16", "bf16"):
        return pb.run16(steps, precision, rounding == "stochastic", want_fields=True)
    return pb.run(steps, "f32" if precision == "fp32" else "f64", want_fields=True)


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16", "bf16"])
@pytest.mark.parametrize("upload", [True, False])
def test_room_without_boundary_list(gpu, precision, upload):
    """Only the domain border bounds the air: the list is empty, uploaded as
    such or never uploaded, and the step is the stencil plus the points."""
    from virs_amd.fdtd import FDTD
    nx, ny, nz = 22, 17, 13
    flags = np.zeros((nz, ny, nx), np.uint8)
    steps = 45
    rng = np.random.default_rng(3)
    air = np.flatnonzero(np.pad(np.ones((nz - 2, ny - 2, nx - 2), bool), 1).reshape(-1))
    src = rng.choice(air, 2, replace=False)
    rcv = rng.choice(air, 3, replace=False)
    sig = np.stack([np.sin(np.arange(steps) * 0.4), np.exp(-np.arange(steps) / 5.0)])
    h = 0.05
    dt = O.dt_for_lambda2(1 / 3, h)
    _say(f"empty list {precision} upload={upload}")
    f = FDTD((nx, ny, nz), h, dt=dt, precision=precision)
    f.upload_flags(flags)
    if upload:
        f.upload_boundary(np.zeros(0, np.uint32))
    f.set_soft_sources(src, sig)
    f.set_receivers(rcv)
    f.start(steps)
    f.step(steps)
    pb = O.Problem(flags, h, dt=dt, soft_idx=src, soft_sig=sig, rcv_idx=rcv)
    ref = _run_oracle(pb, steps, precision)
    np.testing.assert_array_equal(f.receivers(True), ref["ir"])
    np.testing.assert_array_equal(f.field(0), ref["p"])
    f.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_points_at_their_limits(gpu, precision):
    """64 soft sources (distinct cells, 20 on list cells) and 64 receivers
    (20 on list cells, 4 of them sharing cells with others)."""
    from helpers import gpu_context
    p = tiny_problem(13, steps=50, nrcv=4)
    rng = np.random.default_rng(14)
    flags = p["flags"].reshape(-1)
    nz, ny, nx = p["flags"].shape
    inside = np.zeros_like(p["flags"], bool)
    inside[1:nz - 1, 1:ny - 1, 1:nx - 1] = True
    air = np.flatnonzero((flags == 0) & inside.reshape(-1))
    bidx = p["bidx"].astype(np.int64)
    on_list = rng.choice(bidx, 40, replace=False)
    src = np.concatenate([rng.choice(air, 44, replace=False), on_list[:20]])
    rcv = np.concatenate([rng.choice(air, 40, replace=False), on_list[20:40], on_list[20:24]])
    assert len(src) == 64 and len(set(src.tolist())) == 64 and len(rcv) == 64
    sig = rng.standard_normal((64, p["steps"])) * 0.01
    p["soft_idx"], p["soft_sig"], p["rcv_idx"] = src, sig, rcv
    _say(f"points at their limits {precision}")
    f = gpu_context(p, precision)
    f.step(p["steps"])
    ref = _run_oracle(oracle_problem(p), p["steps"], precision)
    ir = f.receivers(True)
    np.testing.assert_array_equal(ir, ref["ir"])
    np.testing.assert_array_equal(ir[60:64], ir[40:44])  # shared cells read alike
    np.testing.assert_array_equal(f.field(0), ref["p"])
    f.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_soft_source_on_a_wall_or_border_is_rejected(gpu, precision):
    """A soft source must sit on a cell that updates: one on a wall or on
    the domain border is rejected when the sources are set after the flags,
    and by the flag upload when they were set before."""
    from virs_amd.fdtd import FDTD, FDTDError
    p = tiny_problem(3)
    flags = p["flags"]
    nz, ny, nx = flags.shape
    wall = int(np.flatnonzero(flags.reshape(-1) == 1)[0])
    border = 0
    air = int(p["soft_idx"][0])
    sig = np.ones((1, 10))
    _say(f"soft source on a wall {precision}")
    f = FDTD((nx, ny, nz), p["h"], dt=p["dt"], precision=precision)
    f.upload_flags(flags)
    for cell in (wall, border):
        with pytest.raises(FDTDError, match="wall or border"):
            f.set_soft_sources([cell], sig)
    f.set_soft_sources([air], sig)
    f.close()
    g = FDTD((nx, ny, nz), p["h"], dt=p["dt"], precision=precision)
    g.set_soft_sources([wall], sig)
    with pytest.raises(FDTDError, match="wall or border"):
        g.upload_flags(flags)
    g.close()
