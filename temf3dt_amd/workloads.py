"""Synthetic models for tests and benchmarks (tensor grids split into Kuhn
tets, square loops along grid lines).

Coordinates follow the reference frame: z positive downward, air at z < 0.
Vertex ids are renumbered so that every transmitter loop's vertices come
first, in path order: each loop is then a chain of MEDIT edges with
ascending vertex ids closed by one descending segment -- the orientation
the reference's source signs assume (+1 for every segment but the last,
src/FETD_FWD.f90:446-448).
"""
import dataclasses
import math

import numpy as np

from . import meshgen

AIR, EARTH = 11, 12


def axis(lo, hi, fine_lo, fine_hi, h, growth):
    """Grid lines: uniform spacing h on [fine_lo, fine_hi], geometric
    growth outward to the domain ends lo, hi (last cell absorbed)."""
    core = list(np.arange(fine_lo, fine_hi + 0.5 * h, h))
    left, x, step = [], fine_lo, h
    while True:
        step *= growth
        x -= step
        if x <= lo + 0.3 * step:
            left.append(lo)
            break
        left.append(x)
    right, x, step = [], fine_hi, h
    while True:
        step *= growth
        x += step
        if x >= hi - 0.3 * step:
            right.append(hi)
            break
        right.append(x)
    return np.array(sorted(left) + core + right, dtype=np.float64)


@dataclasses.dataclass
class Model:
    xyz: np.ndarray
    tet_vex: np.ndarray
    tet_phy: np.ndarray
    loops: list               # list of [nseg,2] int32 arrays (1-based vertex ids)
    receivers: np.ndarray     # [R,3]
    domains: list             # [(phys, sigma, eps_r, mu_r)]
    source: tuple = (1, 1, 1.0e-7, 1.0)   # source_type, pulse_type, t0, current
    sp_division: int = 10
    time_maximum: float = 1e-3
    sp_double: int = 100
    sp_dblesize: int = 2
    time_points: list = None
    rec_charac: float = 10.0
    aniso: list = None        # [(phys, sig3, sig_ang3, per3, per_ang3, mag3, mag_ang3)] or None

    @property
    def ntet(self):
        return len(self.tet_vex)

    def sigma_inv_mu(self):
        sig = np.empty(self.ntet)
        imu = np.empty(self.ntet)
        for ph, s, _e, mr in self.domains:
            m = self.tet_phy == ph
            sig[m] = s
            imu[m] = 1.0 / (mr * (math.pi * 4.0e-7))
        return sig, imu

    def write(self, root, loop=0):
        """ROOT.mesh + ROOT.cntl for transmitter `loop` (reference format)."""
        meshgen.write_medit(root + ".mesh", self.xyz, self.tet_vex, self.tet_phy, self.loops[loop])
        meshgen.write_cntl(root + ".cntl", self.receivers, self.domains, self.source, self.sp_division,
                           self.time_maximum, self.sp_double, self.sp_dblesize, time_points=self.time_points,
                           rec_charac=self.rec_charac, aniso=self.aniso)


def build(xs, ys, zs, loops, receivers, blocks=(), host_sigma=0.01, air_sigma=1e-12, **kw):
    """Tensor grid with an air half-space (z<0), an earth half-space and
    rectangular conductive blocks [(x0,x1,y0,y1,z0,z1,sigma), ...]."""
    xyz, tets, hexid = meshgen.tensor_mesh(xs, ys, zs)
    phy, domains = _materials(xs, ys, zs, hexid, blocks, host_sigma, air_sigma)
    # loop vertex chains in grid-node ids
    chains = []
    for (x0, x1, y0, y1, z) in loops:
        segs = meshgen.square_loop(xs, ys, zs, x0, x1, y0, y1, z)
        chains.append(segs[:, 0].copy())  # path order
    return _loops_first(xyz, tets, phy, chains, domains, receivers, kw)


def _materials(xs, ys, zs, hexid, blocks, host_sigma, air_sigma):
    """Physical group of every tet from the centre of its grid hexahedron
    (hexid [T,3]): air above z = 0, earth below, blocks by containment."""
    cx = 0.5 * (xs[hexid[:, 0]] + xs[hexid[:, 0] + 1])
    cy = 0.5 * (ys[hexid[:, 1]] + ys[hexid[:, 1] + 1])
    cz = 0.5 * (zs[hexid[:, 2]] + zs[hexid[:, 2] + 1])
    phy = np.where(cz < 0, AIR, EARTH).astype(np.int32)
    domains = [(AIR, air_sigma, 1.0, 1.0), (EARTH, host_sigma, 1.0, 1.0)]
    for bi, (x0, x1, y0, y1, z0, z1, s) in enumerate(blocks):
        m = (cx > x0) & (cx < x1) & (cy > y0) & (cy < y1) & (cz > z0) & (cz < z1)
        phy[m] = 13 + bi
        domains.append((13 + bi, s, 1.0, 1.0))
    return phy, domains


def near_wire(loops, r):
    """Predicate (x, y, z) -> True within distance r of the wire of one of
    the square loops [(x0, x1, y0, y1, z)]; loops are looked up in 16 m
    buckets of the horizontal plane."""
    B = 16.0
    buckets = {}
    for lp in loops:
        x0, x1, y0, y1, _ = lp
        for bx in range(math.floor((x0 - r) / B), math.floor((x1 + r) / B) + 1):
            for by in range(math.floor((y0 - r) / B), math.floor((y1 + r) / B) + 1):
                buckets.setdefault((bx, by), []).append(lp)
    r2 = r * r

    def near(x, y, z):
        for x0, x1, y0, y1, zl in buckets.get((math.floor(x / B), math.floor(y / B)), ()):
            dz = z - zl
            dx = max(x0 - x, x - x1, 0.0)
            dy = max(y0 - y, y - y1, 0.0)
            if dx > 0.0 or dy > 0.0:
                dh2 = dx * dx + dy * dy
            else:  # inside the square: distance to its nearest side
                dh = min(x - x0, x1 - x, y - y0, y1 - y)
                dh2 = dh * dh
            if dh2 + dz * dz <= r2:
                return True
        return False
    return near


def build_refined(xs, ys, zs, loops, receivers, blocks=(), host_sigma=0.01, air_sigma=1e-12, h_wire=2.2,
                  r_wire=0.5, margin=3, **kw):
    """build()'s model with the mesh refined at the transmitters: the Kuhn
    tets of the hexahedra within `margin` grid cells of the loops' extent
    are bisected (newest-vertex bisection, which keeps the mesh conforming)
    until no tet with a vertex or its centroid within r_wire of a wire has an
    edge longer than h_wire; the refinement grades back to the grid's cells
    through a few tets.  Materials are those of build() (by the grid
    hexahedron containing each tet); every loop follows its cut grid edges
    as a chain of shorter wire segments."""
    from . import refine
    near = near_wire(loops, r_wire)

    def want(pts, le):
        return le > h_wire and any(near(*q) for q in pts)

    def span(a, v0, v1):
        return (max(int(np.searchsorted(a, v0)) - margin, 0), min(int(np.searchsorted(a, v1)) + margin, len(a) - 1))

    box = (span(xs, min(lp[0] for lp in loops), max(lp[1] for lp in loops)),
           span(ys, min(lp[2] for lp in loops), max(lp[3] for lp in loops)),
           span(zs, min(lp[4] for lp in loops), max(lp[4] for lp in loops)))
    xyz, tets, r, _ = refine.refined_tensor_mesh(xs, ys, zs, box, want)
    cen = xyz[tets].mean(axis=1)
    hexid = np.stack([np.searchsorted(a, cen[:, d], side="right") - 1 for d, a in enumerate((xs, ys, zs))], axis=1)
    phy, domains = _materials(xs, ys, zs, hexid, blocks, host_sigma, air_sigma)
    chains = []
    for (x0, x1, y0, y1, z) in loops:
        path = meshgen.square_loop(xs, ys, zs, x0, x1, y0, y1, z)[:, 0] - 1
        ch = []
        for q in range(len(path)):
            ch += r.expand_edge(int(path[q]), int(path[(q + 1) % len(path)]))[:-1]
        chains.append(np.array(ch, dtype=np.int64) + 1)
    return _loops_first(xyz, tets + 1, phy, chains, domains, receivers, kw)


def _loops_first(xyz, tets, phy, chains, domains, receivers, kw):
    """The model with its vertices renumbered so that every loop's vertices
    come first, in path order (tets and chains 1-based)."""
    N = len(xyz)
    newid = np.zeros(N + 1, np.int64)
    nxt = 1
    for ch in chains:
        for v in ch:
            if newid[v] == 0:
                newid[v] = nxt
                nxt += 1
    rest = np.nonzero(newid[1:] == 0)[0] + 1
    newid[rest] = np.arange(nxt, nxt + len(rest))
    order = np.empty(N, np.int64)
    order[newid[1:] - 1] = np.arange(N)
    xyz_n = xyz[order]
    tets_n = newid[tets].astype(np.int32)
    loop_segs = []
    for ch in chains:
        ids = newid[ch]
        segs = np.stack([ids, np.roll(ids, -1)], axis=1).astype(np.int32)
        loop_segs.append(segs)
    return Model(xyz=xyz_n, tet_vex=tets_n, tet_phy=phy, loops=loop_segs, receivers=np.asarray(receivers, float),
                 domains=domains, **kw)


def small_model():
    """A few-thousand-tet half-space with a 100 m loop (tests, smoke)."""
    xs = axis(-2000, 2000, -100, 100, 50.0, 2.0)
    ys = axis(-2000, 2000, -100, 100, 50.0, 2.0)
    zs = axis(-2000, 2000, -50, 100, 50.0, 2.0)
    return build(xs, ys, zs, loops=[(-50, 50, -50, 50, 0.0)], receivers=[[0.0, 0.0, -3.0], [10.0, 5.0, 7.0]],
                 host_sigma=0.01, air_sigma=1e-8, source=(1, 1, 1.0e-6, 1.0), sp_division=10,
                 time_maximum=1e-3, sp_double=12, sp_dblesize=2, rec_charac=100.0)


def aniso_small_model():
    """small_model with anisotropic media: a tilted uniaxial earth (0.01 S/m
    across, 0.03 S/m along an axis rotated 30/20/10 degrees) and a weakly
    anisotropic, axis-aligned permeability, air isotropic (tests).  The
    permeability stays axis-aligned: the curl weight is its entrywise
    reciprocal, which is its inverse only for a diagonal tensor."""
    m = small_model()
    aniso = []
    for ph, sig, er, mr in m.domains:
        if sig > 1e-6:  # earth
            aniso.append((ph, (0.01, 0.01, 0.03), (30.0, 20.0, 10.0), (er, er, er), (0.0, 0.0, 0.0),
                          (mr, mr * 1.05, mr), (0.0, 0.0, 0.0)))
        else:
            aniso.append((ph, (sig, sig, sig), (0.0, 0.0, 0.0), (er, er, er), (0.0, 0.0, 0.0),
                          (mr, mr, mr), (0.0, 0.0, 0.0)))
    return dataclasses.replace(m, aniso=aniso, time_maximum=2e-4)


def plain_3d_model(scale=1.0):
    """3-D plain large-loop model: 0.01 S/m half-space, 1 S/m block of
    200 x 200 x 100 m, top at 100 m depth, offset 150 m from the centre of
    an 800 m square loop on the ground; 41 receivers every 20 m, 3 m above
    ground; 1e-12 S/m air in a 100 km cube.  About 1.1M tets at scale 1."""
    h = 50.0 / scale
    hz = 20.0 / scale
    xs = axis(-50000, 50000, -450, 450, h, 1.4)
    ys = axis(-50000, 50000, -450, 450, h, 1.4)
    zs = axis(-50000, 50000, -40, 300, hz, 1.4)
    rec = [[-400.0 + 20.0 * i, 0.0, -3.0] for i in range(41)]
    gates = list(np.logspace(-5, -2, 40))
    return build(xs, ys, zs, loops=[(-400, 400, -400, 400, 0.0)], receivers=rec,
                 blocks=[(50, 250, -100, 100, 100, 200, 1.0)], host_sigma=0.01, air_sigma=1e-12,
                 source=(1, 1, 1.0e-6, 1.0), sp_division=10, time_maximum=1.0e-2, sp_double=100, sp_dblesize=2,
                 time_points=gates, rec_charac=60.0)


def multi_source_small_model(nside=2):
    """nside x nside 20 m loops on a 50 m grid over a half-space with a
    block, one coincident receiver 1 m above each loop centre (tests)."""
    cs = [-25.0 * (nside - 1) + 50.0 * i for i in range(nside)]
    xs = axis(-1500, 1500, cs[0] - 20, cs[-1] + 20, 10.0, 2.5)
    ys = axis(-1500, 1500, cs[0] - 20, cs[-1] + 20, 10.0, 2.5)
    zs = axis(-1500, 1500, -20, 40, 10.0, 2.5)
    loops = [(cx - 10, cx + 10, cy - 10, cy + 10, 0.0) for cy in cs for cx in cs]
    rec = [[cx, cy, -1.0] for cy in cs for cx in cs]
    return build(xs, ys, zs, loops=loops, receivers=rec, blocks=[(-30, 30, -30, 30, 20, 50, 0.5)],
                 host_sigma=0.01, air_sigma=1e-8, source=(1, 1, 1.0e-6, 1.0), sp_division=10,
                 time_maximum=5e-5, sp_double=12, sp_dblesize=2, rec_charac=30.0)


def batched_3d_model(nside=8):
    """Multi-transmitter model: the 3-D earth of plain_3d_model (0.01 S/m
    host, 1 S/m block, 1e-12 S/m air) with nside x nside 20 m square loops
    (1 A) on a 50 m grid, each with a coincident receiver 1 m above its
    centre; 10 m cells under every transmitter, ~1.5M tets at nside 8.
    Gates 1e-6..1e-3 s."""
    cs = [-25.0 * (nside - 1) + 50.0 * i for i in range(nside)]
    lo, hi = cs[0] - 20.0, cs[-1] + 20.0
    xs = axis(-50000, 50000, lo, hi, 10.0, 1.75)
    ys = axis(-50000, 50000, lo, hi, 10.0, 1.75)
    zs = axis(-50000, 50000, -20, 200, 10.0, 1.75)
    loops = [(cx - 10, cx + 10, cy - 10, cy + 10, 0.0) for cy in cs for cx in cs]
    rec = [[cx, cy, -1.0] for cy in cs for cx in cs]
    gates = list(np.logspace(-6, -3, 40))
    return build(xs, ys, zs, loops=loops, receivers=rec, blocks=[(50, 250, -100, 100, 100, 200, 1.0)],
                 host_sigma=0.01, air_sigma=1e-12, source=(1, 1, 1.0e-7, 1.0), sp_division=10,
                 time_maximum=1.0e-3, sp_double=100, sp_dblesize=2, time_points=gates, rec_charac=20.0)


def batched_3d_refined_model(nside=8, r_wire=0.5):
    """batched_3d_model's transmitters and earth on a locally refined mesh:
    the Kuhn tets of its 10 m grid that touch a wire (within r_wire) are
    bisected down to 1.25 m cubes (longest edge 2.17 m), grading back to
    10 m within a few cells; each 20 m loop becomes 64 wire segments of
    1.25 m.  1,690,764 tets at nside 8."""
    cs = [-25.0 * (nside - 1) + 50.0 * i for i in range(nside)]
    lo, hi = cs[0] - 20.0, cs[-1] + 20.0
    xs = axis(-50000, 50000, lo, hi, 10.0, 1.75)
    ys = axis(-50000, 50000, lo, hi, 10.0, 1.75)
    zs = axis(-50000, 50000, -20, 200, 10.0, 1.75)
    loops = [(cx - 10, cx + 10, cy - 10, cy + 10, 0.0) for cy in cs for cx in cs]
    rec = [[cx, cy, -1.0] for cy in cs for cx in cs]
    gates = list(np.logspace(-6, -3, 40))
    return build_refined(xs, ys, zs, loops=loops, receivers=rec, blocks=[(50, 250, -100, 100, 100, 200, 1.0)],
                         host_sigma=0.01, air_sigma=1e-12, h_wire=2.2, r_wire=r_wire, source=(1, 1, 1.0e-7, 1.0),
                         sp_division=10, time_maximum=1.0e-3, sp_double=100, sp_dblesize=2, time_points=gates,
                         rec_charac=20.0)


def refined_small_model():
    """small_model's earth and 100 m loop with the tets at the wire bisected
    to 12.5 m cubes (tests)."""
    xs = axis(-2000, 2000, -100, 100, 50.0, 2.0)
    ys = axis(-2000, 2000, -100, 100, 50.0, 2.0)
    zs = axis(-2000, 2000, -50, 100, 50.0, 2.0)
    return build_refined(xs, ys, zs, loops=[(-50, 50, -50, 50, 0.0)], receivers=[[0.0, 0.0, -3.0], [10.0, 5.0, 7.0]],
                         host_sigma=0.01, air_sigma=1e-8, h_wire=22.0, r_wire=2.0, source=(1, 1, 1.0e-6, 1.0),
                         sp_division=10, time_maximum=1e-3, sp_double=12, sp_dblesize=2, rec_charac=100.0)


def uneven_model(target_tets, host_sigma, block, seed_name="", time_maximum=1e-2):
    """One independent model: a half-space with one rectangular block and a
    500 m square loop on the ground (corners at +-250 m), central receiver
    1 m above it; the cell size across the loop and the outward growth
    (1.3-3.0, out to +-50 km) are chosen so that the graded tensor grid has
    about `target_tets` tets."""
    best = None
    for g in (1.3, 1.5, 2.0, 3.0):
        for k in range(2, 80):
            h = 500.0 / k
            xs = axis(-50000, 50000, -250, 250, h, g)
            zs = axis(-50000, 50000, -2 * h, 250, h, g)
            nt = 6 * (len(xs) - 1) ** 2 * (len(zs) - 1)
            if best is None or abs(math.log(nt / target_tets)) < abs(math.log(best[0] / target_tets)):
                best = (nt, h, xs, zs)
    _, h, xs, zs = best
    gates = list(np.logspace(-5, -2, 40))
    return build(xs, xs, zs, loops=[(-250.0, 250.0, -250.0, 250.0, 0.0)], receivers=[[0.0, 0.0, -1.0]],
                 blocks=[block], host_sigma=host_sigma, air_sigma=1e-12, source=(1, 1, 1.0e-6, 1.0),
                 sp_division=10, time_maximum=time_maximum, sp_double=100, sp_dblesize=2, time_points=gates,
                 rec_charac=1.5 * h)


def uneven_models(n=48, seed=3, tmin=2e4, tmax=3e5):
    """Independent models of uneven size: host sigma log-uniform in
    [1e-3, 1e-1] S/m, one block with sigma log-uniform in [1e-2, 10] S/m,
    sides 50-300 m, top 20-300 m deep, tet counts log-uniform in [tmin, tmax]
    (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        host = 10 ** rng.uniform(-3, -1)
        bsig = 10 ** rng.uniform(-2, 1)
        sx, sy, sz = rng.uniform(50, 300, 3)
        top = rng.uniform(20, 300)
        cx, cy = rng.uniform(-150, 150, 2)
        nt = 10 ** rng.uniform(math.log10(tmin), math.log10(tmax))
        out.append(uneven_model(nt, host, (cx - sx / 2, cx + sx / 2, cy - sy / 2, cy + sy / 2, top, top + sz, bsig)))
    return out


def union(models):
    """The models' meshes in one vertex numbering (model blocks in order) for
    one context: xyz, tets, per-tet sigma and 1/mu, per-model wire segments
    (global ids), receivers, and the per-model tet and receiver counts."""
    xyz, tets, sig, imu, loops, recs, ntet, nrec = [], [], [], [], [], [], [], []
    off = 0
    for m in models:
        xyz.append(m.xyz)
        tets.append(m.tet_vex + off)
        s, i = m.sigma_inv_mu()
        sig.append(s)
        imu.append(i)
        loops.append(m.loops[0] + off)
        recs.append(m.receivers)
        ntet.append(m.ntet)
        nrec.append(len(m.receivers))
        off += len(m.xyz)
    return dict(xyz=np.vstack(xyz), tet_vex=np.ascontiguousarray(np.vstack(tets), np.int32), sigma=np.concatenate(sig),
                inv_mu=np.concatenate(imu), loops=loops, receivers=np.vstack(recs), ntet=np.array(ntet),
                nrec=np.array(nrec))


def adjoint_blocks(seed=4):
    """True model of the adjoint workload: three blocks of 1, 0.1 and
    0.003 S/m with sides of 100-300 m (seeded) under the receiver line of
    plain_3d_model, in a 0.01 S/m host."""
    rng = np.random.default_rng(seed)
    out = []
    for s, x in zip((1.0, 0.1, 0.003), (-250.0, 0.0, 250.0)):
        sx, sy, sz = rng.uniform(100, 300, 3)
        top = rng.uniform(40, 150)
        cx = x + rng.uniform(-50, 50)
        cy = rng.uniform(-50, 50)
        out.append((cx - sx / 2, cx + sx / 2, cy - sy / 2, cy + sy / 2, top, top + sz, s))
    return out


def tet_sigma(m, host_sigma, blocks=(), air_sigma=1e-12):
    """Per-tet conductivity of model m's mesh for another earth: host_sigma
    below z = 0, air above, and blocks [(x0,x1,y0,y1,z0,z1,sigma)] by tet
    centroid (later blocks win)."""
    c = m.xyz[m.tet_vex - 1].mean(axis=1)
    sig = np.where(c[:, 2] < 0, air_sigma, host_sigma)
    for (x0, x1, y0, y1, z0, z1, s) in blocks:
        sel = (c[:, 0] > x0) & (c[:, 0] < x1) & (c[:, 1] > y0) & (c[:, 1] < y1) & (c[:, 2] > z0) & (c[:, 2] < z1)
        sig[sel] = s
    return sig
