"""Local refinement of the Kuhn tensor meshes by newest-vertex bisection.

A tensor grid split into Kuhn tetrahedra (six per hexahedron along its main
diagonal, `meshgen.tensor_mesh`) is refined locally by Maubach's tagged
bisection: a tetrahedron (x0, x1, x2, x3) with tag k in {1, 2, 3} is cut at
the midpoint z of its refinement edge (x0, xk) into

    (x0, ..., x_{k-1}, z, x_{k+1}, ..., x3)  and  (x1, ..., x_k, z, x_{k+1}, ..., x3)

both with tag k - 1 (3 after 1).  Every Kuhn tetrahedron, with its vertices
in the order of the cube path from corner 0 to corner 7 and tag 3, starts
with the cube diagonal as refinement edge, and the bisections of a Kuhn
mesh stay conforming and shape-regular: three generations make one uniform
refinement of each cube.  Conformity is kept by the usual recursion: before
an edge is cut, every tetrahedron around it that has another refinement edge
is refined first, then all tetrahedra around the edge are cut at one shared
midpoint.

Only the hexahedra inside a given index box are refined; an edge on a face
of the box that tetrahedra outside it share is never cut (the refinement
raises an error instead, asking for a larger box); faces on the grid's own
boundary may be cut.
"""
import sys

import numpy as np

from . import meshgen

# Kuhn split of the unit cube, each tet a cube path from corner 0 to corner 7
# (corner index = dx + 2 dy + 4 dz): the split meshgen.tensor_mesh makes
KUHN_PATHS = meshgen._KUHN


class Refiner:
    """Kuhn tetrahedra of the hexahedra [i0, i1) x [j0, j1) x [k0, k1) of the
    tensor grid xs x ys x zs (node id (i ny + j) nz + k, 0-based), refined by
    tagged bisection."""

    def __init__(self, xs, ys, zs, box):
        self.xs, self.ys, self.zs = (np.asarray(a, float) for a in (xs, ys, zs))
        self.ny, self.nz = len(ys), len(zs)
        self.box = box
        (i0, i1), (j0, j1), (k0, k1) = box
        # box faces shared with tets outside it (not on the grid's own boundary)
        self.inner_faces = [[lim for lim, end in ((lo, 0), (hi, n - 1)) if lim != end]
                            for (lo, hi), n in zip(box, (len(xs), len(ys), len(zs)))]
        self.coords = []          # coordinates of the midpoints added (ids from the grid's node count on)
        self.nbase = len(xs) * len(ys) * len(zs)
        self.tets = []            # [v0, v1, v2, v3] in bisection order
        self.tag = []
        self.gen = []
        self.alive = []
        self.star = {}            # edge (a, b), a < b -> set of live tet ids
        self.mid = {}             # cut edge -> midpoint id
        for i in range(i0, i1):
            for j in range(j0, j1):
                for k in range(k0, k1):
                    c = [self.nid(i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1)) for q in range(8)]
                    for p in KUHN_PATHS:
                        self._add([c[q] for q in p], 3, 0)

    def nid(self, i, j, k):
        return (i * self.ny + j) * self.nz + k

    def xyz(self, v):
        """Coordinates of vertex v as a tuple of floats."""
        if v < self.nbase:
            i, r = divmod(v, self.ny * self.nz)
            j, k = divmod(r, self.nz)
            return (float(self.xs[i]), float(self.ys[j]), float(self.zs[k]))
        return self.coords[v - self.nbase]

    def _on_box_face(self, a, b):
        """True if the edge a-b lies in a face of the refined box that tets
        outside the box share."""
        if a >= self.nbase or b >= self.nbase:
            pa, pb = self.xyz(a), self.xyz(b)
            axes = (self.xs, self.ys, self.zs)
            return any(pa[d] == pb[d] and any(pa[d] == axes[d][lim] for lim in self.inner_faces[d]) for d in range(3))
        ia, ra = divmod(a, self.ny * self.nz)
        ja, ka = divmod(ra, self.nz)
        ib, rb = divmod(b, self.ny * self.nz)
        jb, kb = divmod(rb, self.nz)
        for (u, w), lims in zip(((ia, ib), (ja, jb), (ka, kb)), self.inner_faces):
            if u == w and u in lims:
                return True
        return False

    def _add(self, v, tag, gen):
        t = len(self.tets)
        self.tets.append(v)
        self.tag.append(tag)
        self.gen.append(gen)
        self.alive.append(True)
        for p in range(4):
            for q in range(p + 1, 4):
                e = (v[p], v[q]) if v[p] < v[q] else (v[q], v[p])
                self.star.setdefault(e, set()).add(t)
        return t

    def _remove(self, t):
        v = self.tets[t]
        self.alive[t] = False
        for p in range(4):
            for q in range(p + 1, 4):
                e = (v[p], v[q]) if v[p] < v[q] else (v[q], v[p])
                s = self.star[e]
                s.discard(t)
                if not s:
                    del self.star[e]

    def refinement_edge(self, t):
        v = self.tets[t]
        a, b = v[0], v[self.tag[t]]
        return (a, b) if a < b else (b, a)

    def _midpoint(self, e):
        m = self.mid.get(e)
        if m is None:
            if self._on_box_face(*e):
                raise ValueError("refinement reaches the boundary of the refined box: enlarge the box")
            m = self.nbase + len(self.coords)
            pa, pb = self.xyz(e[0]), self.xyz(e[1])
            self.coords.append((0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1]), 0.5 * (pa[2] + pb[2])))
            self.mid[e] = m
        return m

    def _bisect(self, t, z):
        v, k, g = self.tets[t], self.tag[t], self.gen[t]
        self._remove(t)
        nk = k - 1 if k > 1 else 3
        self._add(v[:k] + [z] + v[k + 1:], nk, g + 1)
        self._add(v[1:k + 1] + [z] + v[k + 1:], nk, g + 1)

    def refine(self, t):
        """Bisects t, with every tetrahedron around its refinement edge, after
        refining those whose refinement edge differs."""
        e = self.refinement_edge(t)
        while True:
            other = [s for s in self.star[e] if self.refinement_edge(s) != e]
            if not other:
                break
            self.refine(other[0])
        z = self._midpoint(e)
        for s in list(self.star[e]):
            self._bisect(s, z)

    def refine_where(self, want, max_passes=64):
        """Refines until want(t) (given the live tet id) is false for every
        live tetrahedron."""
        old = sys.getrecursionlimit()
        sys.setrecursionlimit(max(old, 100000))
        try:
            for _ in range(max_passes):
                todo = [t for t in range(len(self.tets)) if self.alive[t] and want(t)]
                if not todo:
                    return
                for t in todo:
                    if self.alive[t]:
                        self.refine(t)
            raise RuntimeError("refinement did not finish")
        finally:
            sys.setrecursionlimit(old)

    def live(self):
        return [t for t in range(len(self.tets)) if self.alive[t]]

    def expand_edge(self, a, b):
        """The chain of vertices from a to b along the (possibly cut) edge a-b."""
        e = (a, b) if a < b else (b, a)
        m = self.mid.get(e)
        if m is None:
            return [a, b]
        left = self.expand_edge(a, m)
        return left[:-1] + self.expand_edge(m, b)


def refined_tensor_mesh(xs, ys, zs, box, want_pts):
    """Kuhn mesh of xs x ys x zs with the hexahedra of the index box
    ((i0, i1), (j0, j1), (k0, k1)) refined by bisection while
    want_pts(points, longest_edge) is true, points = the tet's four vertices
    and its centroid (tuples).  Returns xyz [N,3], tets [T,4] (0-based,
    positively oriented), the refiner (for loop chains) and the number of
    tetrahedra in the box."""
    r = Refiner(xs, ys, zs, box)

    def want(t):
        p = [r.xyz(q) for q in r.tets[t]]
        le2 = max((p[a][0] - p[b][0]) ** 2 + (p[a][1] - p[b][1]) ** 2 + (p[a][2] - p[b][2]) ** 2
                  for a in range(4) for b in range(a + 1, 4))
        p.append((0.25 * (p[0][0] + p[1][0] + p[2][0] + p[3][0]), 0.25 * (p[0][1] + p[1][1] + p[2][1] + p[3][1]),
                  0.25 * (p[0][2] + p[1][2] + p[2][2] + p[3][2])))
        return want_pts(p, le2 ** 0.5)

    r.refine_where(want)
    base_xyz, base_tets, hexid = meshgen.tensor_mesh(xs, ys, zs)
    (i0, i1), (j0, j1), (k0, k1) = box
    inside = ((hexid[:, 0] >= i0) & (hexid[:, 0] < i1) & (hexid[:, 1] >= j0) & (hexid[:, 1] < j1)
              & (hexid[:, 2] >= k0) & (hexid[:, 2] < k1))
    outer = base_tets[~inside].astype(np.int64) - 1
    inner = np.array([r.tets[t] for t in r.live()], dtype=np.int64)
    xyz = np.concatenate([base_xyz, np.array(r.coords).reshape(-1, 3)], axis=0)
    tets = np.concatenate([outer, inner], axis=0)
    p = xyz[tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    neg = vol < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return xyz, tets, r, len(inner)


def check_conforming(xyz, tets, lo, hi):
    """Every face is shared by exactly two tetrahedra except faces on the
    boundary of the box lo..hi (one tetrahedron); every tetrahedron has
    positive volume; the volumes add up to the box's.  Raises on failure."""
    p = xyz[tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6.0
    if not np.all(vol > 0):
        raise AssertionError(f"{np.sum(vol <= 0)} tetrahedra without positive volume")
    total = np.prod(np.asarray(hi, float) - np.asarray(lo, float))
    if abs(vol.sum() - total) > 1e-9 * total:
        raise AssertionError(f"volume {vol.sum()} against the box's {total}")
    faces = np.sort(np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]]), axis=1)
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    if np.any(cnt > 2):
        raise AssertionError("a face shared by more than two tetrahedra")
    single = uf[cnt == 1]
    q = xyz[single]
    on_face = np.zeros(len(single), bool)
    for d in range(3):
        for lim in (lo[d], hi[d]):
            on_face |= np.all(q[:, :, d] == lim, axis=1)
    if not np.all(on_face):
        raise AssertionError(f"{np.sum(~on_face)} interior faces with one tetrahedron (hanging vertices)")
