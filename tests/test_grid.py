"""Host grid pipeline (virs_host.cpp) against the numpy restatement in
oracle/voxel_oracle.py and against closed forms.  CPU only."""
import os

import numpy as np
import pytest

from oracle import voxel_oracle as VO
from virs_amd import _lib
from virs_amd import grid as G
from virs_amd import rooms


def random_tris(rng, n, lo, hi):
    return rng.uniform(lo, hi, size=(n, 3, 3)).astype(np.float32)


def test_library_exports_declared_symbols():
    import ctypes
    for name, syms in (("libvirs_host.so", _lib.HOST_SYMBOLS), ("libvirs_hip.so", _lib.HIP_SYMBOLS)):
        lib = ctypes.CDLL(os.path.join(_lib.LIB_DIR, name))
        for s in syms:
            assert hasattr(lib, s), f"{name} lacks {s}"


def test_headers_declare_exactly_the_exported_entry_points():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    import re
    for hdr, syms in (("virs_fdtd.h", _lib.HIP_SYMBOLS), ("virs_host.h", _lib.HOST_SYMBOLS)):
        text = open(os.path.join(root, "include", hdr)).read()
        declared = set(re.findall(r"\b(virs_[a-z0-9_]+)\s*\(", text))
        assert declared == set(syms), (hdr, declared ^ set(syms))


@pytest.mark.parametrize("dims", [(2, 64, 64), (64, 64, 2), (1 << 31, 3, 3), (65536, 65536, 2),
                                  (65536, 65536, 3), (1 << 30, 1 << 30, 1 << 30), (-5, 64, 64)])
def test_create_rejects_grids_outside_the_limits(dims):
    """Dimensions below 3 and grids above 2^32 cells (the boundary list's
    uint32 index range) fail with VIRS_ERR_ARG before any device call, so
    this runs without a GPU.  The 2^90-cell case would overflow int64."""
    import ctypes as C
    L = _lib.hip()
    d = _lib.VirsDesc()
    assert L.virs_desc_default(C.byref(d)) == 0
    d.nx, d.ny, d.nz = dims
    ctx = C.c_void_p(1)
    assert L.virs_create(C.byref(d), C.byref(ctx)) == -1
    assert not ctx.value


@pytest.mark.parametrize("seed", range(6))
def test_voxelize_matches_numpy_restatement(seed):
    rng = np.random.default_rng(seed)
    dims = (21, 17, 15)
    h = 0.1
    origin = np.array([-0.05, 0.02, -0.11], np.float32)
    tri = random_tris(rng, 12, [-0.2, -0.2, -0.2], [2.3, 1.9, 1.7])
    mat = rng.integers(-1, 3, 12).astype(np.int32)
    roles = (0, -2)
    got, _ = G.voxelize(G.Mesh(tri, mat), origin, h, dims, roles)
    want = VO.voxelize_ref(tri, mat, origin, h, dims, roles)
    np.testing.assert_array_equal(got, want)


def test_reference_sat_is_a_superset_of_textbook():
    """The wrong-axis projection (simulator.cpp:73-110) never drops a true
    overlap; it over-fills tilted triangles; x-planes are exact."""
    rng = np.random.default_rng(9)
    dims = (18, 18, 18)
    over = 0
    for _ in range(8):
        tri = random_tris(rng, 1, [1, 1, 1], [16, 16, 16])
        mat = np.zeros(1, np.int32)
        ref = VO.voxelize_ref(tri, mat, np.zeros(3), 1.0, dims)
        txt = VO.voxelize_ref(tri, mat, np.zeros(3), 1.0, dims, test=VO.sat_textbook)
        assert np.all(ref[txt == 1] == 1)
        over += int((ref == 1).sum() - (txt == 1).sum())
    assert over > 0
    # a triangle in the plane x = 7.5 (through voxel centres): exact
    t = np.array([[[7.5, 2.2, 3.1], [7.5, 14.7, 5.9], [7.5, 6.3, 15.2]]], np.float32)
    ref = VO.voxelize_ref(t, np.zeros(1, np.int32), np.zeros(3), 1.0, dims)
    txt = VO.voxelize_ref(t, np.zeros(1, np.int32), np.zeros(3), 1.0, dims, test=VO.sat_textbook)
    np.testing.assert_array_equal(ref, txt)


def test_shoebox_walls_closed_form():
    """The 5x4x3 m shoebox: walls at voxel centres x 17.5/142.5, y 15.5/115.5,
    z 16.5/91.5 give exactly six planes of wall voxels."""
    sc = rooms.shoebox_5x4x3()
    prep = rooms.prepare(sc)
    f = prep.room.flags
    want = np.zeros_like(f)
    xs, ys, zs = slice(17, 143), slice(15, 116), slice(16, 92)
    want[zs, ys, 17] = want[zs, ys, 142] = 1
    want[zs, 15, xs] = want[zs, 115, xs] = 1
    want[16, ys, xs] = want[91, ys, xs] = 1
    walls = (f == 1).astype(np.uint8)
    np.testing.assert_array_equal(walls, want)
    # the shell: face neighbours of walls that are air, inside and outside
    np.testing.assert_array_equal(f, VO.shell_ref(want))
    src = prep.src_idx[0]
    assert f.reshape(-1)[src] == 0


def test_shell_normals_list_match_restatement():
    rng = np.random.default_rng(4)
    f = (rng.random((13, 11, 17)) < 0.2).astype(np.uint8)
    f[rng.random(f.shape) < 0.05] = 3
    sh = G.build_shell(f)
    np.testing.assert_array_equal(sh, VO.shell_ref(f))
    np.testing.assert_array_equal(G.compute_normals(sh), VO.normals_ref(sh))
    np.testing.assert_array_equal(G.boundary_list(sh), VO.boundary_ref(sh))


def test_boundary_materials_rule():
    f = np.zeros((5, 5, 5), np.uint8)
    w = np.zeros_like(f)
    f[2, 2, 3] = 1
    w[2, 2, 3] = 4
    f[2, 3, 2] = 1
    w[2, 3, 2] = 2
    sh = G.build_shell(f)
    idx = G.boundary_list(sh)
    mats = G.boundary_materials(sh, w, idx)
    centre = (2 * 5 + 2) * 5 + 2
    # +x is checked before +y
    assert mats[list(idx).index(centre)] == 4
    assert np.all(G.boundary_materials(sh, None, idx) == 0)


def test_obj_reader_fan_triangulation_and_materials(tmp_path):
    obj = tmp_path / "q.obj"
    (tmp_path / "q.mtl").write_text("newmtl wall\nKd 1 1 1\nnewmtl speaker\nKd 1 0 0\n")
    obj.write_text(
        "mtllib q.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 2\n"
        "usemtl speaker\nf 1/1/1 2/2/2 3/3/3 4/4/4\n"
        "usemtl unknown\nf -1 -2 -3\n"
        "usemtl wall\nf 1 2 5 # comment\n")
    m = G.load_obj(obj)
    assert m.tri.shape == (4, 3, 3)
    np.testing.assert_array_equal(m.mat, [1, 1, -1, 0])
    np.testing.assert_array_equal(m.tri[1], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(m.tri[2], [[0.5, 0.5, 2], [0, 1, 0], [1, 1, 0]])
    np.testing.assert_array_equal(m.vmin, [0, 0, 0])
    np.testing.assert_array_equal(m.vmax, [1, 1, 2])


def test_obj_writer_round_trip(tmp_path):
    sc = rooms.shoebox_5x4x3()
    p = tmp_path / "box.obj"
    rooms.write_obj(str(p), sc.tri, sc.mat)
    m = G.load_obj(p)
    np.testing.assert_array_equal(m.tri, sc.tri)
    np.testing.assert_array_equal(m.mat, sc.mat)


def test_grid_from_bounds_reference_rule():
    """dims = (size_t)(span * (1/h)) in float32 after a 0.1 m nudge."""
    h = G.reference_cell_size()
    assert np.float32(h) == np.float32(343.0) / (np.float32(22050.0) * np.float32(2.0) * np.float32(2.0))
    vmin = np.array([-2.0, 0.0, -3.0], np.float32)
    vmax = np.array([3.0, 3.0, 1.0], np.float32)
    origin, dims = G.grid_from_bounds(vmin, vmax, h)
    hf = np.float32(h)
    for a in range(3):
        lo = vmin[a] - np.float32(0.1)
        hi = vmax[a] + np.float32(0.1)
        assert origin[a] == lo
        assert dims[a] == int((hi - lo) * (np.float32(1) / hf))


def test_bounds_keep_the_flt_min_quirk():
    """bounding_box.max starts at FLT_MIN (simulator.cpp:290): a mesh entirely
    at negative coordinates keeps max = FLT_MIN on that axis."""
    tri = np.array([[[-3, -2, -1], [-2, -3, -1], [-2, -2, -2]]], np.float32)
    m = G.mesh_from_arrays(tri, [0])
    assert m.vmax[0] == np.finfo(np.float32).tiny


def test_rbj_highpass_closed_form():
    out = np.zeros(5, np.float32)
    L = _lib.host()
    fs, fc, q = 48000.0, 440.0, 0.707
    assert L.virs_rbj_highpass(fs, fc, q, 1.0, out.ctypes.data_as(_lib.c_fp)) == 0
    b0, b1, b2, a1, a2 = (float(v) for v in out)
    assert abs(b0 + b1 + b2) < 1e-6                                        # zero at DC
    assert (b0 - b1 + b2) / (1 - a1 + a2) == pytest.approx(1.0, rel=1e-5)  # unity at Nyquist
    w = 2 * np.pi * fc / fs
    z = np.exp(-1j * w)
    H = (b0 + b1 * z + b2 * z * z) / (1 + a1 * z + a2 * z * z)
    assert abs(H) == pytest.approx(q, rel=2e-4)                            # |H(w0)| = Q
    assert L.virs_rbj_lowpass(fs, fc, q, 1.0, out.ctypes.data_as(_lib.c_fp)) == 0
    b0, b1, b2, a1, a2 = (float(v) for v in out)
    # 1 + a1 + a2 is ~3e-3 here, so float32 coefficients leave ~3e-5 of slack
    assert (b0 + b1 + b2) / (1 + a1 + a2) == pytest.approx(1.0, rel=1e-4)  # unity at DC


def test_reference_default_highpass_row():
    """The reference's only material: RBJ high-pass at fs = 1/dt, 440 Hz,
    Q 0.1 (simulator.cpp:955-962)."""
    out = np.zeros(5, np.float32)
    h = np.float32(G.reference_cell_size())
    dt = np.float32(0.5) * h / (np.float32(343) * np.sqrt(np.float32(3)))
    fs = np.float32(1.0 / float(dt))
    _lib.host().virs_rbj_highpass(float(fs), 440.0, 0.1, 1.0, out.ctypes.data_as(_lib.c_fp))
    np.testing.assert_allclose(out, [0.956697, -1.913394, 0.956697, -1.913355, 0.913433], atol=2e-6)


REF_ASSETS = "/root/reference/assets"


@pytest.mark.skipif(not os.path.isdir(REF_ASSETS), reason="reference assets not mounted")
def test_reference_room_assets():
    """room_with_source2.obj (the default config): 968 quads + 96 triangles,
    two materials.  At the reference cell size the float32 rule gives
    1337x900x1079: span_z * (1/h) = 4.2f * 257.14285f = 1079.9999f truncates
    to 1079 (exact arithmetic would give 1080)."""
    m = G.load_obj(os.path.join(REF_ASSETS, "room_with_source2.obj"))
    assert set(np.unique(m.mat)) == {0, 1}
    origin, dims = G.grid_from_bounds(m.vmin, m.vmax, G.reference_cell_size())
    assert dims == (1337, 900, 1079)
    # voxelize at a coarse cell size and check the pipeline invariants
    room = G.build_room(m, h=0.05)
    f = room.flags
    assert (f == 1).any() and (f == 3).any() and (f == 2).any()
    np.testing.assert_array_equal(f, VO.shell_ref(np.where(f == 2, 0, f).astype(np.uint8)))
    np.testing.assert_array_equal(room.bidx, VO.boundary_ref(f))
