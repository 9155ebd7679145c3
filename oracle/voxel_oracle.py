"""numpy restatement of the reference's host grid build -- TEST INFRASTRUCTURE
ONLY (the checker for virs_amd/csrc/host/virs_host.cpp).

Vectorised over voxels, in float32 operations (each numpy float32 op rounds
once, as the C++ does without FP contraction):
  voxelize_ref   simulator.cpp:396-452 with triangleBoxOverlap3 (:45-141),
                 including its edge-axis projection onto (y,z) for all nine tests
  sat_textbook   the textbook separating-axis test (float64), for the
                 superset property the reference's test has
  shell_ref      spongeKernel, sponge_kernel.cu:8-62, literal 26-neighbour scan
  normals_ref    simulator.cpp:476-531
  boundary_ref   simulator.cpp:571-605
"""
import numpy as np

f32 = np.float32


def _sat_ref(g0, g1, g2, bi, bj, bk):
    """Reference triangle/unit-box test for boxes [i,i+1]x[j,j+1]x[k,k+1]
    (arrays bi, bj, bk), float32."""
    bmin = [bi.astype(f32), bj.astype(f32), bk.astype(f32)]
    bmax = [bi.astype(f32) + f32(1), bj.astype(f32) + f32(1), bk.astype(f32) + f32(1)]
    c = [(bmin[a] + bmax[a]) * f32(0.5) for a in range(3)]
    e = [(bmax[a] - bmin[a]) * f32(0.5) for a in range(3)]
    v0 = [f32(g0[a]) - c[a] for a in range(3)]
    v1 = [f32(g1[a]) - c[a] for a in range(3)]
    v2 = [f32(g2[a]) - c[a] for a in range(3)]
    f0 = [v1[a] - v0[a] for a in range(3)]
    f1 = [v2[a] - v1[a] for a in range(3)]
    f2 = [v0[a] - v2[a] for a in range(3)]
    ok = np.ones(bi.shape, bool)

    def axis(a, b, fa, fb, ex, ey):
        p0 = a * v0[1] - b * v0[2]
        p1 = a * v1[1] - b * v1[2]
        p2 = a * v2[1] - b * v2[2]
        mn = np.minimum(np.minimum(p0, p1), p2)
        mx = np.maximum(np.maximum(p0, p1), p2)
        rad = fa * ey + fb * ex
        return ~((mn > rad) | (mx < -rad))

    for f in (f0, f1, f2):
        ax = [np.abs(t) for t in f]
        ok &= axis(f[2], f[1], ax[2], ax[1], e[1], e[2])
        ok &= axis(f[2], f[0], ax[2], ax[0], e[0], e[2])
        ok &= axis(f[1], f[0], ax[1], ax[0], e[0], e[1])
    for a in range(3):
        mn = np.minimum(v0[a], np.minimum(v1[a], v2[a]))
        mx = np.maximum(v0[a], np.maximum(v1[a], v2[a]))
        ok &= ~((mn > e[a]) | (mx < -e[a]))
    n = [f0[1] * f1[2] - f0[2] * f1[1], f0[2] * f1[0] - f0[0] * f1[2], f0[0] * f1[1] - f0[1] * f1[0]]
    d = -((n[0] * v0[0] + n[1] * v0[1]) + n[2] * v0[2])
    r = (e[0] * np.abs(n[0]) + e[1] * np.abs(n[1])) + e[2] * np.abs(n[2])
    ok &= ~(np.abs(d) > r)
    return ok


def sat_textbook(g0, g1, g2, bi, bj, bk):
    """Akenine-Moeller triangle/box overlap with all 13 axes, float64."""
    g = np.array([g0, g1, g2], np.float64)
    c = np.stack([bi + 0.5, bj + 0.5, bk + 0.5], -1)
    v = g[None, :, :] - c[:, None, :]
    e = np.full(3, 0.5)
    edges = [v[:, 1] - v[:, 0], v[:, 2] - v[:, 1], v[:, 0] - v[:, 2]]
    ok = np.ones(len(c), bool)
    axes = []
    for f in edges:
        for u in np.eye(3):
            axes.append(np.cross(np.broadcast_to(u, f.shape), f))
    axes += [np.broadcast_to(u, v[:, 0].shape) for u in np.eye(3)]
    axes.append(np.cross(edges[0], edges[1]))
    for a in axes:
        p = np.einsum("nk,ntk->nt", a, v)
        r = np.abs(a) @ e
        ok &= ~((p.min(1) > r) | (p.max(1) < -r))
    return ok


def voxelize_ref(tri, mat, origin, h, dims, roles=(0, -2), test=_sat_ref):
    """Flags from triangles, last writer wins in triangle order."""
    nx, ny, nz = dims
    flags = np.zeros((nz, ny, nx), np.uint8)
    origin = np.asarray(origin, f32)
    h = f32(h)
    for t, m in zip(np.asarray(tri, f32), mat):
        if m < 0 or m >= len(roles) or roles[m] == -1:
            continue
        g = [(t[v] - origin) / h for v in range(3)]
        mn = np.minimum(np.minimum(g[0], g[1]), g[2])
        mx = np.maximum(np.maximum(g[0], g[1]), g[2])
        lo = [min(max(int(np.floor(mn[a])), 0), dims[a] - 1) for a in range(3)]
        hi = [min(max(int(np.ceil(mx[a])), 0), dims[a] - 1) for a in range(3)]
        kk, jj, ii = np.meshgrid(np.arange(lo[2], hi[2] + 1), np.arange(lo[1], hi[1] + 1),
                                 np.arange(lo[0], hi[0] + 1), indexing="ij")
        ii, jj, kk = ii.reshape(-1), jj.reshape(-1), kk.reshape(-1)
        hit = test(g[0], g[1], g[2], ii, jj, kk)
        flags[kk[hit], jj[hit], ii[hit]] = 3 if roles[m] == -2 else 1
    return flags


def shell_ref(flags):
    """spongeKernel: air cells with a wall at squared distance < N*N+1 = 2
    among their 26 neighbours become 2; border cells keep their input."""
    nz, ny, nx = flags.shape
    out = flags.copy()
    inner = flags[1:-1, 1:-1, 1:-1]
    min_d2 = np.full(inner.shape, 1e30, np.float32)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == 0 and dy == 0 and dz == 0:
                    continue
                nb = flags[1 + dz:nz - 1 + dz, 1 + dy:ny - 1 + dy, 1 + dx:nx - 1 + dx]
                d2 = np.float32(dx * dx + dy * dy + dz * dz)
                min_d2 = np.where(nb == 1, np.minimum(min_d2, d2), min_d2)
    res = np.where(inner != 0, inner, np.where(min_d2 < np.float32(2), 2, 0)).astype(np.uint8)
    out[1:-1, 1:-1, 1:-1] = res
    return out


def normals_ref(flags):
    out = np.zeros_like(flags)
    c = (slice(1, -1),) * 3
    f = flags

    def nb(dz, dy, dx):
        nz, ny, nx = f.shape
        return f[1 + dz:nz - 1 + dz, 1 + dy:ny - 1 + dy, 1 + dx:nx - 1 + dx]

    code = np.zeros(f[c].shape, np.uint8)
    code |= np.where(nb(0, 0, -1) == 0, 2, 0).astype(np.uint8)
    code |= np.where(nb(0, 0, 1) == 0, 1, 0).astype(np.uint8)
    code |= np.where(nb(0, -1, 0) == 0, 8, 0).astype(np.uint8)
    code |= np.where(nb(0, 1, 0) == 0, 4, 0).astype(np.uint8)
    code |= np.where(nb(-1, 0, 0) == 0, 32, 0).astype(np.uint8)
    code |= np.where(nb(1, 0, 0) == 0, 16, 0).astype(np.uint8)
    out[c] = code
    return out


def boundary_ref(flags):
    inner = np.zeros(flags.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    return np.flatnonzero((flags == 2) & inner).astype(np.uint32)
