"""Tensor-grid tetrahedral meshes for the benchmark and test workloads.

The reference meshes its models with gmsh (README.md section 3), which is not
available here.  This generator builds graded tensor grids in the
reference's coordinate frame (z positive downward, air at z < 0) and splits
every hexahedron into six Kuhn tetrahedra sharing the main diagonal, so the
mesh is conforming and every tet has positive signed volume
(src/FETD_UTILITY.f90:105-112 enters the element matrices with its sign).

Loops are square and follow grid lines, written as a chain of MEDIT "Edges"
with increasing vertex order and a closing segment that runs backwards, the
convention the reference's source signs assume (all +1 except the last
segment, src/FETD_FWD.f90:446-448).
"""
import numpy as np

# Kuhn split of the unit cube: 6 tets along the (0,0,0)-(1,1,1) diagonal.
# Vertex index of a cube corner = dx + 2*dy + 4*dz.
_KUHN = [
    (0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7),
]


def tensor_mesh(xs, ys, zs):
    """Nodes and Kuhn tets of the tensor grid xs x ys x zs.  Returns
    xyz [N,3], tet_vex [T,4] (1-based, positive volume), and the (i,j,k)
    index of each tet's hexahedron."""
    nx, ny, nz = len(xs), len(ys), len(zs)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return (i * ny + j) * nz + k + 1

    I, J, K = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    corners = np.stack([nid(I + (c & 1), J + ((c >> 1) & 1), K + ((c >> 2) & 1)) for c in range(8)], axis=1)
    tets = np.concatenate([corners[:, list(t)] for t in _KUHN], axis=0).astype(np.int64)
    hexid = np.concatenate([np.stack([I, J, K], axis=1)] * 6, axis=0)
    # orient every tet positively
    p = xyz[tets - 1]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    neg = vol < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return xyz, tets.astype(np.int32), hexid


def square_loop(xs, ys, zs, x0, x1, y0, y1, z):
    """Vertex chain (1-based ids) of the square loop x0..x1, y0..y1 at depth
    z, along grid lines, as MEDIT segments: ascending vertex ids except the
    closing segment."""
    ny, nz = len(ys), len(zs)

    def idx(a, v):
        i = int(np.argmin(np.abs(a - v)))
        if abs(a[i] - v) > 1e-6 * max(1.0, abs(v)):
            raise ValueError(f"loop coordinate {v} is not a grid line")
        return i

    i0, i1, j0, j1, k = idx(xs, x0), idx(xs, x1), idx(ys, y0), idx(ys, y1), idx(zs, z)

    def nid(i, j):
        return (i * ny + j) * nz + k + 1

    path = []
    for i in range(i0, i1 + 1):
        path.append(nid(i, j0))
    for j in range(j0 + 1, j1 + 1):
        path.append(nid(i1, j))
    for i in range(i1 - 1, i0 - 1, -1):
        path.append(nid(i, j1))
    for j in range(j1 - 1, j0, -1):
        path.append(nid(i0, j))
    segs = [(path[q], path[q + 1]) for q in range(len(path) - 1)] + [(path[-1], path[0])]
    return np.array(segs, dtype=np.int32)


def write_medit(path, xyz, tet_vex, tet_phy, seg_vex=None, seg_phy=51):
    with open(path, "w") as f:
        f.write(" MeshVersionFormatted 2\n Dimension\n 3\n Vertices\n")
        f.write(f" {len(xyz)}\n")
        for p in xyz:
            f.write(f" {p[0]:.17g} {p[1]:.17g} {p[2]:.17g} 0\n")
        if seg_vex is not None and len(seg_vex):
            f.write(f" Edges\n {len(seg_vex)}\n")
            for a, b in seg_vex:
                f.write(f" {a} {b} {seg_phy}\n")
        f.write(f" Tetrahedra\n {len(tet_vex)}\n")
        for t, ph in zip(tet_vex, tet_phy):
            f.write(f" {t[0]} {t[1]} {t[2]} {t[3]} {ph}\n")
        f.write(" End\n")


def write_cntl(path, receivers, domains, source, sp_division, time_maximum, sp_double, sp_dblesize=2,
               output=(0, 0, 0, 0, 0, 1, 0, 0, 0), time_points=None, rec_charac=None, aniso=None):
    """.cntl in the reference grammar (src/FETD_FWD.f90:313-487).  `aniso`:
    [(domain, sig3, sig_ang3, per3, per_ang3, mag3, mag_ang3)] writes an
    'anisotropic:' section instead of 'isotropic:'."""
    with open(path, "w") as f:
        f.write(f"receivers: {len(receivers)}\n")
        for r in receivers:
            f.write(f"{float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
        if rec_charac is not None:
            f.write(f"rec_charac: {float(rec_charac)!r}\n")
        if aniso:
            f.write(f"\nanisotropic: {len(aniso)}\n")
            for d in aniso:
                for q in range(3):
                    vals = list(d[1 + 2 * q]) + list(d[2 + 2 * q])
                    f.write(f"{int(d[0])} " + " ".join(repr(float(v)) for v in vals) + "\n")
        else:
            f.write(f"\nisotropic: {len(domains)}\n")
            for d in domains:
                f.write(f"{int(d[0])} {float(d[1])!r} {float(d[2])!r} {float(d[3])!r}\n")
        f.write("\nboundary condition: 0\n\n")
        f.write("! TYPE 1 - STEPON / TYPE 2 - IMPULSE / TYPE 3 - STEPOFF\n")
        f.write(f"source: 1\n{int(source[0])} {int(source[1])} {float(source[2])!r} {float(source[3])!r}\n\n")
        f.write("sp_mthd: 2\n")
        f.write(f"sp_division: {int(sp_division)}\ntime_maximum: {float(time_maximum)!r}\n")
        f.write(f"sp_double: {sp_double}\nsp_dblesize: {sp_dblesize}\n\n")
        f.write("output:\n" + " ".join(str(v) for v in output) + "\n\nvtkout: 0\n")
        if time_points is not None:
            f.write(f"time_points: {len(time_points)}\n")
            for t in time_points:
                f.write(f"{float(t)!r}\n")
