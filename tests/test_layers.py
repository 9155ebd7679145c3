"""Layer images (VectorSpace::layerToImage, vectorspace.cpp:77-148; the image
loop of simulator.cpp:970-1022): the host library's pixel rule against the
oracle's literal restatement, both pinned by a hand-computed image, and the
PNG writer against an independent decoder (Pillow)."""
import struct
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from virs_amd import layers


def _hand_case():
    """3 x 2 x 3 grid (odd width and height), layer 1; layer 0 holds values
    that would show up if the wrong layer were read."""
    nx, ny, nz = 3, 2, 3
    p = np.full((nz, ny, nx), 0.9, np.float32)
    flags = np.full((nz, ny, nx), 2, np.uint8)
    p[:, 1, :] = np.array([[0.1, -0.2, 0.001],
                           [1e-4, -1e-4, 0.5],
                           [-0.0, 0.0999, -0.0015]], np.float32)
    flags[:, 1, :] = np.array([[1, 2, 0],
                               [3, 0, 1],
                               [2, 2, 0]], np.uint8)
    # (r, g, b) by hand: R = trunc(clamp(1500 p)) for p >= 0 (-0.0 counts as
    # >= 0), else B = trunc(clamp(-1500 p)); G = 100 for a wall, 255 for a
    # boundary cell.  0.0999 -> 149.85 -> 149 (truncation, not rounding),
    # -0.2 -> 300 -> 255 (clamp), 0.001 -> 1.5 -> 1.
    cells = [(150, 100, 0), (0, 255, 255), (1, 0, 0),
             (0, 0, 0), (0, 0, 0), (255, 100, 0),
             (0, 255, 0), (149, 255, 0), (0, 0, 2)]
    # the image is 4 x 4; pixel (x, z) sits at flat pixel z*3 + x (the
    # unpadded width), so the nine cells fill flat pixels 0..8 and the rest
    # stays black
    img = np.zeros((4 * 4, 3), np.uint8)
    img[:9] = cells
    return p, flags, img.reshape(4, 4, 3)


def test_layer_image_hand_computed():
    p, flags, want = _hand_case()
    np.testing.assert_array_equal(O.layer_image(p, flags, 1), want)
    np.testing.assert_array_equal(layers.layer_rgb(p[:, 1, :], flags, 1), want)


@pytest.mark.parametrize("dims", [(6, 3, 4), (5, 2, 7), (7, 4, 5), (8, 3, 8), (1, 1, 1), (9, 5, 2)])
def test_layer_rgb_matches_oracle(dims):
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz)
    p = (rng.standard_normal((nz, ny, nx)) * 0.1).astype(np.float32)
    p.reshape(-1)[:: 7] = 0.0
    p.reshape(-1)[3:: 11] = -0.0
    flags = rng.integers(0, 4, (nz, ny, nx)).astype(np.uint8)
    for layer in sorted({0, ny // 2, ny - 1}):
        want = O.layer_image(p, flags, layer)
        got = layers.layer_rgb(p[:, layer, :], flags, layer)
        assert got.shape == (nz + nz % 2, nx + nx % 2, 3)
        np.testing.assert_array_equal(got, want)


def test_layer_out_of_range():
    p, flags, _ = _hand_case()
    with pytest.raises(IndexError):
        O.layer_image(p, flags, 2)
    with pytest.raises(RuntimeError):
        layers.layer_rgb(p[:, 0, :], flags, 2)
    with pytest.raises(RuntimeError):
        layers.layer_rgb(p[:, 0, :], flags, -1)


def test_frame_names():
    """simulator.cpp:990-1013 pads the frame number to six digits."""
    assert layers.frame_name(0, 10) == "output_step_000000.png"
    assert layers.frame_name(70, 10) == "output_step_000007.png"
    assert layers.frame_name(1239, 10) == "output_step_000123.png"
    assert layers.frame_name(123456, 1) == "output_step_123456.png"
    assert layers.frame_name(1234567, 1) == "output_step_1234567.png"


def _chunks(raw):
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    at, out = 8, []
    while at < len(raw):
        n, = struct.unpack(">I", raw[at:at + 4])
        kind = raw[at + 4:at + 8]
        data = raw[at + 8:at + 8 + n]
        crc, = struct.unpack(">I", raw[at + 8 + n:at + 12 + n])
        assert crc == zlib.crc32(kind + data)
        out.append((kind, data))
        at += 12 + n
    return out


@pytest.mark.parametrize("wh", [(4, 4), (7, 3), (512, 256), (1, 1)])
def test_png_round_trip(tmp_path, wh):
    from PIL import Image
    w, h = wh
    rng = np.random.default_rng(w * h)
    rgb = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
    path = tmp_path / "layer.png"
    layers.write_png(path, rgb)
    chunks = _chunks(path.read_bytes())
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    assert chunks[0][1] == struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    with Image.open(path) as im:
        assert im.mode == "RGB" and im.size == (w, h)
        np.testing.assert_array_equal(np.asarray(im), rgb)


def test_png_of_a_sheared_layer(tmp_path):
    from PIL import Image
    p, flags, want = _hand_case()
    path = tmp_path / "output_layer.png"
    layers.write_png(path, layers.layer_rgb(p[:, 1, :], flags, 1))
    with Image.open(path) as im:
        np.testing.assert_array_equal(np.asarray(im), want)
