"""Writes tests/golden/random_512x384x256_ir.npz: the impulse responses of the
512x384x256 random room (6 seeded boxes + 4 tilted slabs, 3 biquads, the
Gaussian-derivative source, 8 receivers) over its full 20000 steps, from the
oracle's reference-literal FP32 run: float filter state and float
coefficients, as simulator_gpu.cu:105-150 keeps them, with the dense update
of simulator_gpu.cu:38-103.

    python tests/golden/make_random_room_ir.py [--threads N] [--work DIR]

The run goes in chunks of 2000 steps that carry the fields, the filter state
and the step offset (pinned bit for bit against one uninterrupted run by
tests/test_oracle.py::test_chunked_runs_equal_one_run); each finished chunk
is saved under --work so an interrupted run resumes.  The fixture also holds
digests of the room's flags and boundary list and its source and receiver
cells, so the test can check that it rebuilt the same room.
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(ROOT, "tests", "golden", "random_512x384x256_ir.npz")
STEPS = 20000
CHUNK = 2000


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--work", default="/tmp/virs_random_room_fixture")
    args = ap.parse_args()
    from oracle import oracle as O
    from virs_amd import rooms

    prep = rooms.prepare(rooms.random_room())
    room, sc = prep.room, prep.scenario
    O.set_threads(args.threads)
    pb = O.Problem(room.flags, sc.h, dt=prep.dt, bidx=room.bidx, bmat=room.bmat, coeffs=sc.coeffs,
                   state_f32=True, soft_idx=prep.src_idx, soft_sig=prep.src_sig, rcv_idx=prep.rcv_idx)
    pb.state_io = np.zeros(4 * max(1, len(room.bidx)), np.float64)
    os.makedirs(args.work, exist_ok=True)
    irs = []
    for n0 in range(0, STEPS, CHUNK):
        path = os.path.join(args.work, f"chunk_{n0:05d}.npz")
        if os.path.exists(path):
            z = np.load(path)
            irs.append(z["ir"])
            pb.init_p, pb.init_q = z["p"], z["q"]
            pb.state_io[:] = z["state"]
            continue
        t0 = time.perf_counter()
        out = pb.run(CHUNK, "f32", n0=n0, want_fields=True)
        pb.init_p = np.ascontiguousarray(out["p"].reshape(-1), np.float64)
        pb.init_q = np.ascontiguousarray(out["q"].reshape(-1), np.float64)
        irs.append(out["ir"])
        np.savez(path + ".tmp.npz", ir=out["ir"], p=pb.init_p, q=pb.init_q, state=pb.state_io)
        os.replace(path + ".tmp.npz", path)
        print(f"steps {n0}..{n0 + CHUNK} in {time.perf_counter() - t0:.0f}s", flush=True)
    ir = np.concatenate(irs, axis=1)
    assert ir.shape == (len(prep.rcv_idx), STEPS) and np.isfinite(ir).all()
    # FP32 run: every sample is a float
    assert np.array_equal(ir.astype(np.float32).astype(np.float64), ir)
    np.savez_compressed(
        OUT, ir=ir.astype(np.float32), steps=STEPS, dims=np.array(room.dims), dt=np.float32(prep.dt),
        coeffs=sc.coeffs, src_idx=prep.src_idx, rcv_idx=prep.rcv_idx, nb=len(room.bidx),
        flags_sha256=digest(room.flags), bidx_sha256=digest(room.bidx), bmat_sha256=digest(room.bmat),
        src_sig_sha256=digest(prep.src_sig))
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
