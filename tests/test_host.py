"""Host-side checks (CPU only): the C ABI library's exports, the symbolic
analysis (ordering, supernodal structure, maps), the mesh generator, and
that the product path fails loudly without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "temf3dt_amd", "libtemfetd_host.so")


def _P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_abi_library_exports_every_declared_symbol():
    from temf3dt_amd import fetd
    lib = fetd.load_library()
    hdr = open(os.path.join(ROOT, "include", "temfetd.h")).read()
    names = sorted(set(re.findall(r"\b(tf_[a-z_0-9]+)\s*\(", hdr)))
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_product_path_fails_loudly_without_gpu():
    import subprocess
    import sys
    code = ("import os; os.environ['HIP_VISIBLE_DEVICES']='-1'\n"
            "from temf3dt_amd import fetd\n"
            "try:\n    fetd.Fetd(0)\nexcept fetd.FetdError as e:\n    print('raised', e)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert "raised" in r.stdout, r.stdout + r.stderr


def _host():
    return ctypes.CDLL(HOST_LIB)


def _set_mesh(H, xyz, tv):
    xyz = np.ascontiguousarray(xyz, np.float64)
    tv = np.ascontiguousarray(tv, np.int32)
    nE, nnz = ctypes.c_int(), ctypes.c_longlong()
    assert H.tfh_set_mesh(len(xyz), _P(xyz), len(tv), _P(tv), ctypes.byref(nE), ctypes.byref(nnz)) == 0
    n, nz = nE.value, nnz.value
    ia = np.zeros(n + 1, np.int32)
    ja = np.zeros(nz, np.int32)
    H.tfh_get_topology(None, None, None, _P(ia), _P(ja))
    return n, nz, ia - 1, ja - 1


def _analyze(H, n, nz, leaf):
    v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_longlong(),
         ctypes.c_longlong(), ctypes.c_longlong()]
    assert H.tfh_analyze(leaf, *[ctypes.byref(x) for x in v]) == 0
    nsn, ns = v[0].value, v[6].value
    perm = np.zeros(n, np.int32)
    c0, nc, par, lev = [np.zeros(nsn, np.int32) for _ in range(4)]
    sptr = np.zeros(nsn + 1, np.int64)
    sidx = np.zeros(ns, np.int32)
    rel = np.zeros(ns, np.int32)
    amap = np.zeros(nz, np.int64)
    loff = np.zeros(nsn + 1, np.int64)
    H.tfh_get_symbolic(_P(perm), _P(c0), _P(nc), _P(par), _P(lev), _P(sptr), _P(sidx), _P(rel), _P(amap), _P(loff))
    return dict(nsn=nsn, nnzL=v[2].value, perm=perm, c0=c0, nc=nc, par=par, lev=lev, sptr=sptr, sidx=sidx,
                rel=rel, amap=amap, loff=loff)


def _two_meshes():
    """Two disjoint meshes of different sizes in one vertex numbering."""
    from temf3dt_amd import meshgen
    xa, ta, _ = meshgen.tensor_mesh(np.linspace(0, 5, 6), np.linspace(0, 4, 5), np.linspace(0, 3, 4))
    xb, tb, _ = meshgen.tensor_mesh(np.linspace(0, 3, 4), np.linspace(0, 2, 3), np.linspace(0, 4, 5))
    return np.vstack([xa, xb]), np.vstack([ta, tb + len(xa)]), len(xa)


@pytest.mark.parametrize("leaf,kind", [(4, "one"), (16, "one"), (4, "two"), (16, "two")])
def test_supernodal_structure_covers_cholesky_fill(leaf, kind):
    """Every nonzero of the dense Cholesky factor of the permuted matrix lies
    in the supernodal structure; relative indices and the assembly map
    decode back to the right (row, column).  For two disjoint meshes the
    fronts form one tree per mesh and no front mixes the two."""
    from temf3dt_amd import meshgen
    if kind == "one":
        xyz, tv, _ = meshgen.tensor_mesh(np.linspace(0, 5, 6), np.linspace(0, 4, 5), np.linspace(0, 6, 7))
    else:
        xyz, tv, nA = _two_meshes()
    H = _host()
    n, nz, ia, ja = _set_mesh(H, xyz, tv)
    s = _analyze(H, n, nz, leaf)
    perm = s["perm"]
    assert sorted(perm.tolist()) == list(range(n))
    iperm = np.empty(n, int)
    iperm[perm] = np.arange(n)
    rng = np.random.default_rng(0)
    A = np.zeros((n, n))
    for r in range(n):
        for q in range(ia[r], ia[r + 1]):
            A[r, ja[q]] = A[ja[q], r] = rng.uniform(-1, 1)
    A += np.eye(n) * (np.abs(A).sum(1) + 1)
    Lc = np.linalg.cholesky(A[np.ix_(perm, perm)])
    allowed = np.zeros((n, n), bool)
    rows_of = []
    for k in range(s["nsn"]):
        cols = np.arange(s["c0"][k], s["c0"][k] + s["nc"][k])
        rows = np.concatenate([cols, s["sidx"][s["sptr"][k]:s["sptr"][k + 1]]])
        rows_of.append(rows)
        for j in cols:
            allowed[rows[rows >= j], j] = True
    assert not np.any((np.abs(Lc) > 1e-14) & ~allowed)
    for k in range(s["nsn"]):
        p = s["par"][k]
        if p >= 0:
            q0, q1 = s["sptr"][k], s["sptr"][k + 1]
            np.testing.assert_array_equal(rows_of[p][s["rel"][q0:q1]], s["sidx"][q0:q1])
            assert s["lev"][p] > s["lev"][k]
    sn_of = np.empty(n, int)
    for k in range(s["nsn"]):
        sn_of[s["c0"][k]:s["c0"][k] + s["nc"][k]] = k
    for r in range(n):
        for q in range(ia[r], ia[r + 1]):
            a, b = iperm[r], iperm[ja[q]]
            lo, hi = min(a, b), max(a, b)
            k = sn_of[lo]
            f = len(rows_of[k])
            ld = (f + 15) // 16 * 16  # columns padded to 16 doubles
            off = s["amap"][q] - s["loff"][k]
            assert s["loff"][k] % 16 == 0
            assert off // ld == lo - s["c0"][k] and off % ld < f and rows_of[k][off % ld] == hi
    if kind == "two":
        edg = np.zeros((n, 2), np.int32)
        H.tfh_get_topology(_P(edg), None, None, None, None)
        in_b = edg[:, 0] > nA  # 1-based vertex ids; mesh B's edges
        assert (s["par"] < 0).sum() == 2
        for k in range(s["nsn"]):
            cols = perm[s["c0"][k]:s["c0"][k] + s["nc"][k]]
            assert in_b[cols].all() or not in_b[cols].any()


def test_generated_mesh_is_conforming_and_positive():
    from temf3dt_amd import workloads
    m = workloads.small_model()
    p = m.xyz[m.tet_vex - 1]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    assert vol.min() > 0
    faces = np.sort(np.concatenate([m.tet_vex[:, [0, 1, 2]], m.tet_vex[:, [0, 1, 3]], m.tet_vex[:, [0, 2, 3]],
                                    m.tet_vex[:, [1, 2, 3]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2  # every face shared by at most two tets
    # loop chain: ascending vertex ids except the closing segment (reference sign convention)
    segs = m.loops[0]
    assert np.all(segs[:-1, 0] < segs[:-1, 1]) and segs[-1, 0] > segs[-1, 1]
    assert np.all(segs[1:, 0] == segs[:-1, 1]) and segs[-1, 1] == segs[0, 0]


def test_plain_3d_workload_size():
    """The 3-D plain model is the ~1M-tet workload of the benchmark."""
    from temf3dt_amd import workloads
    m = workloads.plain_3d_model()
    assert 0.85e6 < m.ntet < 1.2e6
    assert len(m.receivers) == 41 and len(m.loops) == 1
    assert set(np.unique(m.tet_phy)) == {11, 12, 13}


def test_front_size_guard_of_32_bit_kernel_addressing():
    """Fronts whose L block (nc columns of stride f rounded up to 16, plus
    the 64 columns a tile may address past it) or row-major L21 copy reach
    4 GiB are rejected by the analysis; the kernels address both with 32-bit
    byte offsets."""
    H = _host()
    H.tfh_front_fits.argtypes = [ctypes.c_longlong, ctypes.c_longlong]
    # the largest front of the 1.35M-tet batched mesh (10,167 columns) fits
    assert H.tfh_front_fits(10167, 3000) == 1
    # 23,170 columns: (23,170 + 64) * 23,184 * 8 B = 4.309e9 B >= 2^32
    assert H.tfh_front_fits(23170, 0) == 0
    assert H.tfh_front_fits(23104, 0) == 1  # (23,104 + 64) * 23,104 * 8 = 4.282e9 < 2^32
    # a narrow front with a long struct: the L block fits, the copy does not
    assert H.tfh_front_fits(16, 16 * 1024 * 1024 * 2) == 0
    ld = lambda f: (f + 15) // 16 * 16  # noqa: E731
    for nc, m in [(1, 0), (64, 64), (5000, 9000), (20000, 5000), (1000, 500000)]:
        want = (nc + 64) * ld(nc + m) * 8 < 2 ** 32 and m * ld(nc) * 8 < 2 ** 32
        assert H.tfh_front_fits(nc, m) == int(want), (nc, m)
