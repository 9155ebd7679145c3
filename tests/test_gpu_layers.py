"""Layer snapshots on the GPU (the field read of layerToImage,
vectorspace.cpp:77-148, simulator.cpp:970-1022) through the C ABI: planes
bit-exact with the whole-field readback in every storage type, snapshots
queued behind later steps, the ticket rules, and the image loop's PNGs
against the oracle's images of the oracle's fields."""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, tiny_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16", "bf16"])
def test_plane_matches_field_readback(gpu, precision):
    p = tiny_problem(1, dims=(27, 14, 11))
    _say(f"plane snapshot {precision} dims={p['dims']}")
    f = gpu_context(p, precision)
    f.step(37)
    full = f.field(0).astype(np.float32)
    nz, ny, nx = full.shape
    for y in (0, ny // 2, ny - 1):
        np.testing.assert_array_equal(f.plane_fetch(f.plane_request(y)), full[:, y, :])
    f.close()


def test_snapshots_are_queued_behind_their_steps(gpu):
    """A request takes the field as of the steps launched before it, even
    when more steps are launched before the fetch; fetches may come in any
    order.  FP32 against the oracle, bit-exact."""
    p = tiny_problem(0)
    _say("queued snapshots fp32")
    f = gpu_context(p, "fp32", use_graph=True)
    y = p["dims"][1] // 2 + 1
    tickets = []
    for n in (5, 31, 12):
        f.step(n)
        tickets.append(f.plane_request(y))
    planes = [f.plane_fetch(t) for t in reversed(tickets)][::-1]
    ref = oracle_problem(p)
    for s, plane in zip((5, 36, 48), planes):
        want = ref.run(s, "f32", want_fields=True)["p"]
        np.testing.assert_array_equal(plane, want[:, y, :].astype(np.float32))
    f.close()


def test_ticket_rules(gpu):
    p = tiny_problem(2)
    f = gpu_context(p, "fp32")
    f.step(3)
    ny = p["dims"][1]
    t = [f.plane_request(1) for _ in range(4)]
    with pytest.raises(RuntimeError, match="outstanding"):
        f.plane_request(1)
    f.plane_fetch(t[1])
    with pytest.raises(RuntimeError, match="outstanding"):
        f.plane_request(1)  # the next ticket's slot is t[0]'s
    f.plane_fetch(t[0])
    t.append(f.plane_request(2))
    with pytest.raises(RuntimeError, match="no outstanding"):
        f.plane_fetch(t[0])
    with pytest.raises(RuntimeError, match="out of range"):
        f.plane_request(ny)
    with pytest.raises(RuntimeError, match="out of range"):
        f.plane_request(-1)
    for k in (2, 3, 4):
        f.plane_fetch(t[k])
    f.close()


def _png(path):
    from PIL import Image
    with Image.open(path) as im:
        return np.asarray(im).copy()


def test_image_loop_matches_oracle(gpu, tmp_path):
    """simulate's loop with layer images every 10 steps over 35 steps: the
    start image and frames 0..3 (after loop steps 0, 10, 20, 30)."""
    from virs_amd import layers
    p = tiny_problem(1)
    _say("image loop fp32, 35 steps, interval 10")
    f = gpu_context(p, "fp32", use_graph=True)
    y = p["dims"][1] // 2
    first = tmp_path / "output_layer.png"
    paths = layers.simulate_with_layer_images(f, p["flags"], 35, y, interval=10, out_dir=str(tmp_path),
                                              first_image=str(first))
    names = [str(first)] + [str(tmp_path / f"output_step_{k:06d}.png") for k in range(4)]
    assert paths == names
    assert f.steps_done == 35
    ref = oracle_problem(p)
    zero = np.zeros(p["flags"].shape, np.float32)
    np.testing.assert_array_equal(_png(first), O.layer_image(zero, p["flags"], y))
    for k, s in enumerate((1, 11, 21, 31)):
        fld = ref.run(s, "f32", want_fields=True)["p"].astype(np.float32)
        np.testing.assert_array_equal(_png(names[k + 1]), O.layer_image(fld, p["flags"], y))
    f.close()


def test_image_loop_stops_with_the_rms_rule(gpu, tmp_path):
    """With the RMS stop at loop step 3001 and an interval of 3001 the loop
    writes frame 0 and frame 1 (the stopping step's image), then stops."""
    from virs_amd import grid as G
    from virs_amd import layers
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
    stop = O.rms_continue_sequence(pb.run(steps, "f32")["ir"][0].astype(np.float32))
    assert stop == 3001
    _say("image loop with the RMS stop, interval 3001")
    f = FDTD((nx, ny, nz), h, use_graph=True)
    f.upload_flags(flags)
    f.upload_boundary(bidx)
    f.set_materials(row)
    f.set_receivers([ridx])
    f.set_early_stop(True)
    y = ny // 2 + 2
    paths = layers.simulate_with_layer_images(f, flags, steps, y, interval=stop, out_dir=str(tmp_path))
    assert paths == [str(tmp_path / "output_step_000000.png"), str(tmp_path / "output_step_000001.png")]
    assert f.steps_done == stop + 1
    for path, s in zip(paths, (1, stop + 1)):
        fld = pb.run(s, "f32", want_fields=True)["p"].astype(np.float32)
        np.testing.assert_array_equal(_png(path), O.layer_image(fld, flags, y))
    f.close()
