"""Shared builders for the parity tests: small seeded grids that exercise
every cell type, the boundary list with several materials, hard and soft
sources, receivers on boundary and non-boundary cells, and the sponge."""
import numpy as np

from virs_amd import grid as G
from virs_amd import rooms
from oracle import oracle as O


def tiny_problem(seed=0, dims=(24, 18, 14), nmat=2, sponge=3, hard=True, nrcv=4, lam2=1 / 3,
                 h=0.05, steps=60):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    flags = np.zeros((nz, ny, nx), np.uint8)
    wall = np.zeros_like(flags)
    # an inner box of walls plus scattered wall voxels and a thin partition
    flags[2, 2:ny - 2, 2:nx - 2] = 1
    flags[nz - 3, 2:ny - 2, 2:nx - 2] = 1
    flags[2:nz - 2, 2, 2:nx - 2] = 1
    flags[2:nz - 2, ny - 3, 2:nx - 2] = 1
    flags[2:nz - 2, 2:ny - 2, 2] = 1
    flags[2:nz - 2, 2:ny - 2, nx - 3] = 1
    k = rng.integers(4, nz - 4)
    flags[k, 4:ny // 2, 5:nx - 5] = 1
    pts = rng.integers([1, 1, 1], [nz - 1, ny - 1, nx - 1], size=(25, 3))
    for z, y, x in pts:
        flags[z, y, x] = 1
    wall[flags == 1] = rng.integers(0, nmat, int((flags == 1).sum()))
    if hard:
        flags[nz // 2, ny // 2, nx // 2 + 3] = 3
    flags = G.build_shell(flags)
    bidx = G.boundary_list(flags)
    bmat = G.boundary_materials(flags, wall, bidx)
    coeffs = np.array([rooms.random_biquad(rng) for _ in range(nmat)])
    inside = np.zeros_like(flags, bool)
    inside[3:nz - 3, 3:ny - 3, 3:nx - 3] = True
    air = np.flatnonzero((flags == 0) & inside)
    src = rng.choice(air, 2, replace=False)
    # the second soft source sits on a boundary-list cell
    src[1] = bidx[(2 * len(bidx)) // 3]
    rcv = list(rng.choice(air, max(nrcv - 2, 0), replace=False))
    rcv += [int(bidx[len(bidx) // 3]), int(src[0])]
    sig = np.stack([rooms.gauss_deriv(steps, 2.0, 10.0), rooms.unit_impulse(steps) * 0.5])
    dt = O.dt_for_lambda2(lam2, h)
    return dict(flags=flags, bidx=bidx, bmat=bmat, coeffs=coeffs, soft_idx=np.asarray(src, np.int64),
                soft_sig=sig, rcv_idx=np.asarray(rcv[:nrcv], np.int64), h=h, dt=dt, sponge=sponge,
                sigma=0.5, steps=steps, dims=dims)


def oracle_problem(p, state_f32=False, with_sponge=True):
    return O.Problem(p["flags"], p["h"], dt=p["dt"], bidx=p["bidx"], bmat=p["bmat"],
                     coeffs=p["coeffs"], state_f32=state_f32, soft_idx=p["soft_idx"],
                     soft_sig=p["soft_sig"], rcv_idx=p["rcv_idx"],
                     sponge_width=p["sponge"] if with_sponge else 0, sponge_sigma=p["sigma"])


def gpu_context(p, precision="fp32", use_graph=False, dense_blocks=0, fuse=False, rounding="stochastic"):
    from virs_amd.fdtd import FDTD
    f = FDTD(p["dims"], p["h"], dt=p["dt"], precision=precision, sponge_width=p["sponge"],
             sponge_sigma=p["sigma"], use_graph=use_graph, dense_blocks=dense_blocks,
             fuse_boundary=fuse, rounding=rounding)
    f.upload_flags(p["flags"])
    f.upload_boundary(p["bidx"], p["bmat"])
    f.set_materials(p["coeffs"])
    f.set_soft_sources(p["soft_idx"], p["soft_sig"])
    f.set_receivers(p["rcv_idx"])
    f.start(p["steps"])
    return f


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
