"""Local mesh refinement by newest-vertex bisection (temf3dt_amd/refine.py)
and the refined workloads built on it (CPU only).

The bisection is pinned by what the mathematics fixes: three generations of
Maubach bisection of a Kuhn cube are the Kuhn split of its eight half-size
cubes (the tensor mesh of the halved grid, vertex sets compared exactly);
any refinement keeps the mesh conforming (every interior face shared by two
tets, volumes positive and summing to the domain's).  The conformity check
itself is shown to catch a hanging vertex, a flipped tet and a missing tet
planted into a good mesh.
"""
import numpy as np
import pytest

from temf3dt_amd import meshgen, refine, workloads


def _box_mesh(n, want):
    xs = ys = zs = np.arange(n + 1, dtype=float)
    box = ((0, n), (0, n), (0, n))
    xyz, tets, r, cnt = refine.refined_tensor_mesh(xs, ys, zs, box, want)
    return xyz, tets, r, cnt


def _vertex_sets(xyz, tets):
    return sorted(tuple(sorted(tuple(xyz[v]) for v in t)) for t in tets)


def test_three_bisections_are_the_reflected_kuhn_split():
    # every tet bisected three generations: each 0.5 subcube of a unit cube
    # is split into the six Kuhn tets around the diagonal that joins the unit
    # cube's centre to its corner in that subcube (the reflected Kuhn split)
    xyz, tets, r, cnt = _box_mesh(2, lambda pts, le: le > 0.9)
    assert cnt == 8 * 6 * 8
    assert set(r.gen[t] for t in r.live()) == {3}
    refine.check_conforming(xyz, tets, (0, 0, 0), (2, 2, 2))
    p = xyz[tets]
    sub = np.floor(p.mean(axis=1) / 0.5).astype(int)
    groups = {}
    for t, key in enumerate(map(tuple, sub)):
        groups.setdefault(key, []).append(t)
    assert len(groups) == 64 and all(len(g) == 6 for g in groups.values())
    for key, g in groups.items():
        lo = np.array(key) * 0.5
        centre = np.floor(lo) + 0.5
        corner = np.where(lo == centre, np.floor(lo) + 1.0, np.floor(lo))
        vsets = []
        for t in g:
            q = p[t]
            assert any(np.array_equal(v, centre) for v in q) and any(np.array_equal(v, corner) for v in q)
            # a monotone path centre -> corner: the other two vertices differ
            # from the centre in one and two coordinates
            nd = sorted(int(np.count_nonzero(v != centre)) for v in q)
            assert nd == [0, 1, 2, 3]
            vsets.append(tuple(sorted(map(tuple, q))))
        assert len(set(vsets)) == 6


def test_local_refinement_conforms_and_grades():
    # refine near one edge of the box only until edges are at most 0.3 (the
    # generation-8 tets of 1/4 cubes, longest edge 0.25): conforming, graded
    def want(pts, le):
        return le > 0.3 and any(q[0] < 0.6 and q[1] < 0.6 for q in pts)
    xyz, tets, r, cnt = _box_mesh(4, want)
    refine.check_conforming(xyz, tets, (0, 0, 0), (4, 4, 4))
    gens = [r.gen[t] for t in r.live()]
    assert max(gens) == 8 and min(gens) == 0
    p = xyz[tets]
    le = np.sqrt(np.max(np.sum((p[:, :, None] - p[:, None]) ** 2, axis=-1), axis=(1, 2)))
    near = np.any((p[:, :, 0] < 0.6) & (p[:, :, 1] < 0.6), axis=1)
    assert np.all(le[near] <= 0.3)
    # the shape stays in the Kuhn classes: volume ratios are powers of 2
    vol = np.abs(np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])))
    lv = np.log2(vol / vol.max())
    assert np.allclose(lv, np.round(lv))


def test_conformity_check_catches_planted_defects():
    xyz, tets, r, cnt = _box_mesh(2, lambda pts, le: le > 1.0 and min(q[2] for q in pts) < 0.5)
    refine.check_conforming(xyz, tets, (0, 0, 0), (2, 2, 2))
    # a hanging vertex: one tet bisected without the others around its edge
    r2 = refine.Refiner(*(np.arange(3, dtype=float),) * 3, ((0, 2), (0, 2), (0, 2)))
    t = next(t for t in r2.live() if len(r2.star[r2.refinement_edge(t)]) > 1)
    r2._bisect(t, r2._midpoint(r2.refinement_edge(t)))
    live = np.array([r2.tets[q] for q in r2.live()], dtype=np.int64)
    bx = np.concatenate([meshgen.tensor_mesh(*(np.arange(3, dtype=float),) * 3)[0], np.array(r2.coords)])
    p = bx[live]
    neg = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) < 0
    live[neg, 2], live[neg, 3] = live[neg, 3].copy(), live[neg, 2].copy()
    with pytest.raises(AssertionError, match="hanging"):
        refine.check_conforming(bx, live, (0, 0, 0), (2, 2, 2))
    flipped = tets.copy()
    flipped[5, 2], flipped[5, 3] = flipped[5, 3], flipped[5, 2]
    with pytest.raises(AssertionError, match="positive volume"):
        refine.check_conforming(xyz, flipped, (0, 0, 0), (2, 2, 2))
    with pytest.raises(AssertionError, match="volume"):
        refine.check_conforming(xyz, tets[1:], (0, 0, 0), (2, 2, 2))


def test_refinement_stops_at_the_box_boundary():
    xs = np.arange(4, dtype=float)
    with pytest.raises(ValueError, match="boundary of the refined box"):
        refine.refined_tensor_mesh(xs, xs, xs, ((1, 3), (1, 3), (1, 3)), lambda pts, le: le > 0.5)


def test_refined_small_model():
    m = workloads.refined_small_model()
    s = workloads.small_model()
    assert m.ntet == 17592 and s.ntet == 15288
    refine.check_conforming(m.xyz, m.tet_vex - 1, m.xyz.min(0), m.xyz.max(0))
    # same earth: the conductive volume per physical group is unchanged
    for ph in (11, 12):
        def volume(mm):
            p = mm.xyz[mm.tet_vex[mm.tet_phy == ph] - 1]
            return np.sum(np.abs(np.einsum("ij,ij->i", p[:, 1] - p[:, 0],
                                           np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])))) / 6
        assert volume(m) == pytest.approx(volume(s), rel=1e-12)
    # the loop: 32 wire segments of 12.5 m along the 100 m square, ids ascending but the
    # closing segment, every segment an edge of the mesh
    segs = m.loops[0]
    assert len(segs) == 32
    assert np.all(segs[:-1, 0] < segs[:-1, 1]) and segs[-1, 0] > segs[-1, 1]
    assert np.all(segs[:-1, 1] == segs[1:, 0]) and segs[-1, 1] == segs[0, 0]
    d = m.xyz[segs[:, 1] - 1] - m.xyz[segs[:, 0] - 1]
    assert np.allclose(np.linalg.norm(d, axis=1), 12.5)
    assert np.all(np.count_nonzero(d, axis=1) == 1) and np.all(m.xyz[segs[:, 0] - 1, 2] == 0)
    tv = m.tet_vex
    edges = {tuple(sorted((int(t[a]), int(t[b])))) for t in tv for a in range(4) for b in range(a + 1, 4)}
    assert all(tuple(sorted((int(a), int(b)))) in edges for a, b in segs)


def test_near_wire_distance():
    near = workloads.near_wire([(-10, 10, -10, 10, 0.0), (40, 60, -10, 10, 0.0)], 0.5)
    assert near(10.4, 0.0, 0.0) and near(9.6, 3.0, 0.2) and near(40.0, -10.3, 0.3)
    assert not near(0.0, 0.0, 0.0) and not near(10.0, 0.0, 0.6) and not near(25.0, 0.0, 0.0)
    assert near(10.3, 10.3, 0.0) and not near(10.4, 10.4, 0.0)
