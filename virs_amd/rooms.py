"""Synthetic rooms for the benchmark and parity workloads.

Each builder returns a Scenario: triangles + mesh-material ids, the grid
(cell size, origin, dims), Courant number, biquad rows, sources, receivers
and step count.  prepare() runs the host grid pipeline on it and resolves
source/receiver positions to voxel indices.

Workloads (sizes from the benchmark definitions):
  shoebox_5x4x3      5.0 x 4.0 x 3.0 m shoebox, 12 triangles, dx 0.04 m,
                     160 x 132 x 108 cells (128 x 100 x 76 + 16-cell sponge),
                     one material b=[0.6,-0.2,0.05] a=[1,-0.3,0.1], unit
                     impulse at (1.2,1.5,1.1), receiver (3.6,2.4,1.4), 2000 steps
  random_512x384x256 union of 6 seeded boxes + 4 tilted slabs, 3 random stable
                     biquads, Gaussian-derivative source (sigma 4 steps),
                     8 receivers, 20000 steps
  shoebox_grid       the shoebox builder scaled to any grid with M materials
                     (256 x 192 x 128 for the adjoint, 1536 x 1024 x 768 large room)
"""
from dataclasses import dataclass, field
import os

import numpy as np

from . import grid as G


@dataclass
class Scenario:
    name: str
    tri: np.ndarray                 # float32 [T,3,3]
    mat: np.ndarray                 # int32 [T]
    roles: tuple                    # mesh material id -> wall row (>=0) / -2 hard source / -1 ignore
    h: float
    origin: np.ndarray
    dims: tuple
    lam2: float                     # Courant number squared
    coeffs: np.ndarray              # float64 [M,5] b0 b1 b2 a1 a2
    sources: list                   # [(position xyz, signal float64[...])]
    receivers: list                 # [position xyz]
    steps: int
    sponge_width: int = 0
    sponge_sigma: float = 0.0
    air_box: tuple = None           # (lo, hi) metres: where seeded points may go
    meta: dict = field(default_factory=dict)


@dataclass
class Prepared:
    scenario: Scenario
    room: G.Room
    src_idx: np.ndarray             # int64 [S]
    src_sig: np.ndarray             # float64 [S, steps]
    rcv_idx: np.ndarray             # int64 [R]
    dt: np.float32


def box_triangles(lo, hi, face_mats):
    """12 triangles of an axis-aligned box; face_mats per face in the order
    -x, +x, -y, +y, -z, +z."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        [(x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)],
        [(x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)],
        [(x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)],
        [(x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)],
        [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)],
        [(x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)],
    ]
    tri, mat = [], []
    for f, m in zip(faces, face_mats):
        tri += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
        mat += [m, m]
    return np.asarray(tri, np.float32), np.asarray(mat, np.int32)


def write_obj(path, tri, mat, names=None):
    """Writes an OBJ + MTL pair (one `usemtl` group per material run)."""
    base = os.path.splitext(path)[0]
    mids = sorted(set(int(m) for m in mat if m >= 0))
    names = names or {m: f"mat{m}" for m in mids}
    with open(base + ".mtl", "w") as f:
        for m in range(max(mids) + 1 if mids else 0):
            f.write(f"newmtl {names.get(m, f'mat{m}')}\nKd 0.8 0.8 0.8\n\n")
    with open(path, "w") as f:
        f.write(f"mtllib {os.path.basename(base)}.mtl\n")
        for t in tri:
            for v in t:
                f.write("v %.9g %.9g %.9g\n" % tuple(float(c) for c in v))
        cur = None
        for i, m in enumerate(mat):
            if m != cur:
                f.write(f"usemtl {names.get(int(m), 'none')}\n")
                cur = m
            f.write(f"f {3 * i + 1} {3 * i + 2} {3 * i + 3}\n")


def gauss_deriv(steps, sigma=4.0, center=None):
    """Gaussian-derivative pulse, peak magnitude 1, centred at 6 sigma."""
    n = np.arange(steps, dtype=np.float64)
    c = 6.0 * sigma if center is None else center
    s = -(n - c) / sigma * np.exp(-0.5 * ((n - c) / sigma) ** 2)
    return s / np.max(np.abs(s))


def unit_impulse(steps):
    s = np.zeros(steps)
    s[0] = 1.0
    return s


def random_biquad(rng, gain=0.95, npts=4096):
    """Stable biquad: conjugate pole pair at radius U[0.2,0.9], angle U[0,pi],
    zero pair likewise at radius U[0,0.9]; b scaled so max |H| = gain."""
    rp, tp = rng.uniform(0.2, 0.9), rng.uniform(0.0, np.pi)
    rz, tz = rng.uniform(0.0, 0.9), rng.uniform(0.0, np.pi)
    a = np.array([1.0, -2.0 * rp * np.cos(tp), rp * rp])
    b = np.array([1.0, -2.0 * rz * np.cos(tz), rz * rz])
    w = np.linspace(0.0, np.pi, npts)
    z = np.exp(-1j * w)
    H = (b[0] + b[1] * z + b[2] * z * z) / (a[0] + a[1] * z + a[2] * z * z)
    b = b * (gain / np.max(np.abs(H)))
    return np.array([b[0], b[1], b[2], a[1], a[2]])


def shoebox_5x4x3(receivers=1, seed=1234):
    """The 5 x 4 x 3 m shoebox on a 160 x 132 x 108 grid at dx = 0.04 m.
    The origin puts every wall at a voxel centre (x 17.5/142.5, y 15.5/115.5,
    z 16.5/91.5 in voxel units) with 16+ cells of sponge outside the room."""
    tri, mat = box_triangles((0.0, 0.0, 0.0), (5.0, 4.0, 3.0), [0] * 6)
    h = 0.04
    origin = np.array([-0.70, -0.62, -0.66], np.float32)
    rcv = [(3.6, 2.4, 1.4)]
    if receivers > 1:
        rng = np.random.default_rng(seed)
        while len(rcv) < receivers:
            rcv.append(tuple(rng.uniform([0.3, 0.3, 0.3], [4.7, 3.7, 2.7])))
    return Scenario(
        name="shoebox_5x4x3", tri=tri, mat=mat, roles=(0,), h=h, origin=origin,
        dims=(160, 132, 108), lam2=1.0 / 3.0,
        coeffs=np.array([[0.6, -0.2, 0.05, -0.3, 0.1]]),
        sources=[((1.2, 1.5, 1.1), unit_impulse(2000))], receivers=rcv, steps=2000,
        sponge_width=16, sponge_sigma=0.5, air_box=((0.0, 0.0, 0.0), (5.0, 4.0, 3.0)))


def shoebox_grid(dims, nmat=3, h=0.04, steps=4096, nrcv=4, seed=7, lam2=1.0 / 3.0,
                 margin=2, sponge_width=0, sponge_sigma=0.5):
    """Shoebox filling a grid of `dims` cells: walls at voxel centres `margin`
    cells in from the edge; the six faces take materials 0..nmat-1 in turn
    (-x, +x, -y, +y, -z, +z); seeded biquads, Gaussian-derivative source and
    receivers in the air."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    lo = np.array([margin + 0.5] * 3) * h
    hi = (np.array([nx, ny, nz]) - margin - 0.5) * h
    tri, mat = box_triangles(tuple(lo), tuple(hi), [k % nmat for k in range(6)])
    coeffs = np.array([random_biquad(rng) for _ in range(nmat)])
    span = hi - lo
    pts = [tuple(lo + span * rng.uniform(0.15, 0.85, 3)) for _ in range(nrcv + 1)]
    return Scenario(
        name=f"shoebox_{nx}x{ny}x{nz}", tri=tri, mat=mat, roles=tuple(range(nmat)), h=h,
        origin=np.zeros(3, np.float32), dims=(nx, ny, nz), lam2=lam2, coeffs=coeffs,
        sources=[(pts[0], gauss_deriv(steps))], receivers=pts[1:], steps=steps,
        sponge_width=sponge_width, sponge_sigma=sponge_sigma, air_box=(tuple(lo), tuple(hi)))


def _union_boundary(boxes, mats):
    """Triangles of the boundary of a union of axis-aligned boxes, found on
    the grid of their distinct coordinates; each face takes the material of
    the first box that holds the inside cell."""
    xs = sorted(set(v for b in boxes for v in (b[0][0], b[1][0])))
    ys = sorted(set(v for b in boxes for v in (b[0][1], b[1][1])))
    zs = sorted(set(v for b in boxes for v in (b[0][2], b[1][2])))
    cx = [(xs[i] + xs[i + 1]) / 2 for i in range(len(xs) - 1)]
    cy = [(ys[i] + ys[i + 1]) / 2 for i in range(len(ys) - 1)]
    cz = [(zs[i] + zs[i + 1]) / 2 for i in range(len(zs) - 1)]
    owner = -np.ones((len(cx), len(cy), len(cz)), np.int64)
    for bi in reversed(range(len(boxes))):
        lo, hi = boxes[bi]
        for i, x in enumerate(cx):
            for j, y in enumerate(cy):
                for k, z in enumerate(cz):
                    if lo[0] < x < hi[0] and lo[1] < y < hi[1] and lo[2] < z < hi[2]:
                        owner[i, j, k] = bi
    tri, mat = [], []

    def own(i, j, k):
        if 0 <= i < len(cx) and 0 <= j < len(cy) and 0 <= k < len(cz):
            return owner[i, j, k]
        return -1

    for axis, coords in ((0, xs), (1, ys), (2, zs)):
        dims = [len(cx), len(cy), len(cz)]
        rng_ = [range(dims[0]), range(dims[1]), range(dims[2])]
        rng_[axis] = range(dims[axis] + 1)
        for a in rng_[0]:
            for b in rng_[1]:
                for c in rng_[2]:
                    ijk = [a, b, c]
                    inner = own(*ijk)
                    ijk[axis] -= 1
                    outer = own(*ijk)
                    if (inner >= 0) == (outer >= 0):
                        continue
                    m = mats[inner if inner >= 0 else outer]
                    p = coords[[a, b, c][axis]]
                    u_ax, v_ax = [d for d in range(3) if d != axis]
                    grids = (xs, ys, zs)
                    iu, iv = [a, b, c][u_ax], [a, b, c][v_ax]
                    u0, u1 = grids[u_ax][iu], grids[u_ax][iu + 1]
                    v0, v1 = grids[v_ax][iv], grids[v_ax][iv + 1]
                    quad = []
                    for (u, v) in ((u0, v0), (u1, v0), (u1, v1), (u0, v1)):
                        q = [0.0, 0.0, 0.0]
                        q[axis], q[u_ax], q[v_ax] = p, u, v
                        quad.append(q)
                    tri += [[quad[0], quad[1], quad[2]], [quad[0], quad[2], quad[3]]]
                    mat += [m, m]
    return np.asarray(tri, np.float32).reshape(-1, 3, 3), np.asarray(mat, np.int32)


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_room(dims=(512, 384, 256), h=0.02, seed=2024, nmat=3, steps=20000, nrcv=8,
                lam2=1.0 / 3.0):
    """Union of 6 seeded axis-aligned boxes (box 0 the main hall, 1-5 alcoves
    overlapping it) plus 4 tilted slabs inside the hall.  Box faces are snapped
    to voxel centres."""
    rng = np.random.default_rng(seed)
    ext = np.array(dims, np.float64) * h
    snap = lambda v: (np.floor(v / h) + 0.5) * h
    lo0 = snap(ext * 0.2)
    hi0 = snap(ext * 0.8)
    boxes = [(lo0, hi0)]
    for _ in range(5):
        size = ext * rng.uniform(0.2, 0.45, 3)
        c = rng.uniform(lo0, hi0)
        lo = np.clip(snap(c - size / 2), snap(ext * 0.04), None)
        hi = np.clip(snap(c + size / 2), None, snap(ext * 0.96))
        boxes.append((lo, hi))
    box_mats = [int(m) for m in rng.integers(0, nmat, len(boxes))]
    tri, mat = _union_boundary(boxes, box_mats)
    slabs_t, slabs_m = [], []
    for _ in range(4):
        size = np.array([rng.uniform(0.15, 0.35) * ext[0], rng.uniform(0.15, 0.35) * ext[1], 3 * h])
        c = rng.uniform(lo0 + 0.2 * (hi0 - lo0), hi0 - 0.2 * (hi0 - lo0))
        R = _rotation(rng)
        t, _ = box_triangles(tuple(-size / 2), tuple(size / 2), [0] * 6)
        t = (t.reshape(-1, 3).astype(np.float64) @ R.T + c).reshape(-1, 3, 3)
        slabs_t.append(t.astype(np.float32))
        slabs_m.append(np.full(len(t), int(rng.integers(0, nmat)), np.int32))
    tri = np.concatenate([tri] + slabs_t)
    mat = np.concatenate([mat] + slabs_m)
    coeffs = np.array([random_biquad(rng) for _ in range(nmat)])
    pts_rng = np.random.default_rng(seed + 1)
    return Scenario(
        name=f"random_{dims[0]}x{dims[1]}x{dims[2]}", tri=tri, mat=mat, roles=tuple(range(nmat)),
        h=h, origin=np.zeros(3, np.float32), dims=tuple(dims), lam2=lam2, coeffs=coeffs,
        sources=[(None, gauss_deriv(steps))], receivers=[None] * nrcv, steps=steps,
        air_box=(tuple(lo0 + 4 * h), tuple(hi0 - 4 * h)), meta={"points_rng": pts_rng})


def _air_point(room, lo, hi, rng, taken):
    for _ in range(100000):
        p = rng.uniform(lo, hi)
        i = room.index_of(p)
        if room.flags.reshape(-1)[i] == 0 and i not in taken:
            return tuple(p), i
    raise RuntimeError("no air cell found")


def prepare(sc, nthreads=0):
    """Runs the host grid pipeline and resolves sources / receivers."""
    mesh = G.mesh_from_arrays(sc.tri, sc.mat)
    room = G.build_room(mesh, h=sc.h, origin=sc.origin, dims=sc.dims, roles=sc.roles,
                        nthreads=nthreads)
    rng = sc.meta.get("points_rng")
    taken = set()
    src_idx, src_sig = [], []
    for pos, sig in sc.sources:
        if pos is None:
            pos, i = _air_point(room, *sc.air_box, rng, taken)
        else:
            i = room.index_of(pos)
        taken.add(i)
        src_idx.append(i)
        s = np.zeros(sc.steps)
        s[: min(len(sig), sc.steps)] = sig[: sc.steps]
        src_sig.append(s)
    rcv_idx = []
    for pos in sc.receivers:
        if pos is None:
            pos, i = _air_point(room, *sc.air_box, rng, taken)
        else:
            i = room.index_of(pos)
        rcv_idx.append(i)
    dt = np.float32(np.sqrt(sc.lam2) * float(np.float32(sc.h)) / 343.0)
    return Prepared(sc, room, np.asarray(src_idx, np.int64), np.asarray(src_sig, np.float64),
                    np.asarray(rcv_idx, np.int64), dt)


def crop(prep, dims):
    """Corner crop of a prepared room (flags, boundary list and points
    re-derived on the crop): a CPU-sized sample of the same workload."""
    nx, ny, nz = dims
    flags = np.ascontiguousarray(prep.room.flags[:nz, :ny, :nx])
    fl = flags.copy()
    fl[0, :, :] = fl[-1, :, :] = 0
    bidx = G.boundary_list(flags)
    wm = np.ascontiguousarray(prep.room.wall_mat[:nz, :ny, :nx])
    bmat = G.boundary_materials(flags, wm, bidx)
    return flags, bidx, bmat
