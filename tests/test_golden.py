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
