"""The committed fixtures describe the rooms the tests rebuild (no GPU)."""
import hashlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "random_512x384x256_ir.npz")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_random_room_fixture_describes_this_room():
    """tests/golden/random_512x384x256_ir.npz was written for the room that
    rooms.random_room() builds today: same flags, list, materials, source
    signal, source and receiver cells, coefficients and time step."""
    from virs_amd import rooms
    prep = rooms.prepare(rooms.random_room())
    z = np.load(FIXTURE)
    room = prep.room
    assert tuple(z["dims"]) == tuple(room.dims)
    assert str(z["flags_sha256"]) == _digest(room.flags)
    assert str(z["bidx_sha256"]) == _digest(room.bidx)
    assert str(z["bmat_sha256"]) == _digest(room.bmat)
    assert str(z["src_sig_sha256"]) == _digest(prep.src_sig)
    np.testing.assert_array_equal(z["src_idx"], prep.src_idx)
    np.testing.assert_array_equal(z["rcv_idx"], prep.rcv_idx)
    np.testing.assert_array_equal(z["coeffs"], prep.scenario.coeffs)
    assert np.float32(z["dt"]) == prep.dt
    ir = z["ir"]
    assert ir.dtype == np.float32 and ir.shape == (len(prep.rcv_idx), int(z["steps"]))
    assert np.isfinite(ir).all() and (np.abs(ir).max(axis=1) > 0).all()


REF_ROOM = os.path.join(ROOT, "tests", "golden", "room_with_source2.npz")


def test_reference_room_fixture_rebuilds_at_reference_defaults():
    """The shipped room (tests/golden/room_with_source2.npz, written by
    make_reference_room.py from assets/room_with_source2.obj) built by the
    host pipeline at the reference's cell size and bounds rule: the grid
    1337 x 900 x 1079, the list length and the digests of flags, normals,
    list and materials the fixture recorded; and the count of voxels that
    the other quad diagonal changes (the exposure to tinyobjloader's
    unpinned quad rule, simulator.cpp:279)."""
    from virs_amd import grid as G
    z = np.load(REF_ROOM)
    mesh = G.mesh_from_arrays(z["tri"], z["mat"])
    room = G.build_room(mesh)
    assert tuple(room.dims) == (1337, 900, 1079) == tuple(z["dims"])
    assert np.float32(room.h) == z["h"]
    np.testing.assert_array_equal(room.origin, z["origin"])
    assert len(room.bidx) == int(z["nb"])
    np.testing.assert_array_equal(np.bincount(room.flags.reshape(-1), minlength=4), z["counts"])
    assert _digest(room.flags) == str(z["flags_sha256"])
    assert _digest(room.normals) == str(z["normals_sha256"])
    assert _digest(room.bidx) == str(z["bidx_sha256"])
    assert _digest(room.bmat) == str(z["bmat_sha256"])
    pre = G.voxelize(mesh, room.origin, room.h, room.dims)[0]
    alt = G.voxelize(G.mesh_from_arrays(z["alt_tri"], z["mat"]), room.origin, room.h, room.dims)[0]
    assert int(np.count_nonzero(pre != alt)) == int(z["diag_changed_pre_shell"])
