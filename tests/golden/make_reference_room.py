"""Writes tests/golden/room_with_source2.npz: the shipped default room of the
reference (assets/room_with_source2.obj, 968 quads + 96 triangles, materials
0 wall and 1 source) as this library's OBJ reader parses it, with what the
host grid pipeline builds from it at the reference's defaults
(h = 343/88200, the bounds + nudge rule of simulator.cpp:289-329):

    tri, mat            the triangles (quads fan-split v0 v1 v2 / v0 v2 v3)
    dims, origin, h     the grid: 1337 x 900 x 1079
    nb, counts          list length and the count of each flag value
    *_sha256            digests of the flags (after the shell), normals,
                        boundary list and list materials
    alt_tri             the same quads split along the other diagonal
                        (v0 v1 v3 / v1 v2 v3)
    diag_changed_*      voxels whose flag differs between the two splits,
                        before and after the shell: the exposure to the
                        unpinned quad rule of tinyobjloader
                        (simulator.cpp:279; the submodule version is not
                        recorded)

    python tests/golden/make_reference_room.py

Reads /root/reference/assets at generation time only; the tests use the
arrays.  The OBJ's face lines are parsed here as well (positions only), to
build the other split and to check that the two parsers agree.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

SRC = "/root/reference/assets/room_with_source2.obj"
OUT = os.path.join(ROOT, "tests", "golden", "room_with_source2.npz")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def parse_faces(path):
    """Faces as vertex-position index lists with the material in effect
    (material ids in order of first use), plus the positions."""
    pos, faces, mats, names = [], [], [], {}
    cur = -1
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            pos.append([float(np.float32(x)) for x in t[1:4]])
        elif t[0] == "usemtl":
            cur = names.setdefault(t[1], len(names))
        elif t[0] == "f":
            idx = [int(w.split("/")[0]) for w in t[1:]]
            faces.append([i - 1 if i > 0 else len(pos) + i for i in idx])
            mats.append(cur)
    return np.asarray(pos, np.float32), faces, mats


def split(pos, faces, mats, other_diagonal):
    tri, mat = [], []
    for f, m in zip(faces, mats):
        if len(f) == 4 and other_diagonal:
            parts = [(f[0], f[1], f[3]), (f[1], f[2], f[3])]
        else:
            parts = [(f[0], f[k], f[k + 1]) for k in range(1, len(f) - 1)]
        for p in parts:
            tri.append(pos[list(p)])
            mat.append(m)
    return np.asarray(tri, np.float32), np.asarray(mat, np.int32)


def main():
    from virs_amd import grid as G
    mesh = G.load_obj(SRC)
    pos, faces, mats = parse_faces(SRC)
    nquads = sum(len(f) == 4 for f in faces)
    tri_fan, mat_fan = split(pos, faces, mats, False)
    assert np.array_equal(tri_fan, mesh.tri) and np.array_equal(mat_fan, mesh.mat), "parsers disagree"
    tri_alt, mat_alt = split(pos, faces, mats, True)
    room = G.build_room(mesh)
    h = G.reference_cell_size()
    alt = G.build_room(G.mesh_from_arrays(tri_alt, mat_alt), h=h, origin=room.origin, dims=room.dims)
    pre, pre_alt = G.voxelize(mesh, room.origin, h, room.dims)[0], G.voxelize(
        G.mesh_from_arrays(tri_alt, mat_alt), room.origin, h, room.dims)[0]
    f = room.flags
    np.savez_compressed(
        OUT, tri=mesh.tri, mat=mesh.mat, alt_tri=tri_alt, h=np.float32(h), dims=np.array(room.dims),
        origin=room.origin, nb=len(room.bidx), n_quads=nquads,
        counts=np.bincount(f.reshape(-1), minlength=4),
        flags_sha256=digest(f), normals_sha256=digest(room.normals), bidx_sha256=digest(room.bidx),
        bmat_sha256=digest(room.bmat),
        diag_changed_pre_shell=int(np.count_nonzero(pre != pre_alt)),
        diag_changed_after_shell=int(np.count_nonzero(f != alt.flags)),
        diag_nb_alt=len(alt.bidx))
    print(f"dims {room.dims} nb {len(room.bidx)} quads {nquads}; other diagonal changes "
          f"{np.count_nonzero(pre != pre_alt)} voxels before the shell, {np.count_nonzero(f != alt.flags)} after "
          f"(list {len(alt.bidx)})", flush=True)


if __name__ == "__main__":
    main()
