"""The oracle (CPU restatement) pinned against the reference's shipped
outputs, closed forms and brute force.  CPU only."""
import math
import os

import numpy as np
import pytest
from scipy.special import erf

from oracle import pyoracle

UNIT = (np.array([0.0, 1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0, 0.0]), np.array([0.0, 0.0, 0.0, 1.0]))
IND_TET = [(0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3)]  # src/FETD_FWD.f90:23


def test_oracle_reproduces_shipped_field(src_mod_big, tmp_path):
    """The full adaptive run of src/mod_big_sig_0.001 writes a .field that is
    byte-identical to the reference's shipped file (579 rows, every dt change)."""
    pyoracle.set_threads(os.cpu_count() or 8)
    o = pyoracle.Oracle(src_mod_big["mesh"], src_mod_big["cntl"])
    o.assemble()
    out = str(tmp_path / "o.field")
    r = o.run(field=out)
    assert open(out).read() == open(src_mod_big["ref_field"]).read()
    assert len(r["step"]) == 579 and r["nfact"] == 49


def test_oracle_rows_fixture_matches_shipped_field(src_mod_big):
    """The committed full-precision rows print to the shipped 5-digit file."""
    from temf3dt_amd import io as tfio
    g = np.load(src_mod_big["oracle_rows"])
    lines = [tfio.format_row(int(s), float(t), f, [0, 0, 0, 0, 0, 1, 0, 0, 0])
             for s, t, f in zip(g["step"], g["t"], g["fields"])]
    assert "\n".join(lines) + "\n" == open(src_mod_big["ref_field"]).read()


def test_edge_numbering_and_pattern_counts(oracle_mod_big):
    """Invariants of the shipped src mesh: E and nnz of the upper CSR, and
    Euler's formula V - E + F - T = 1 for a ball."""
    o = oracle_mod_big
    assert (o.nE, o.nnz) == (96712, 841453)
    topo = o.topology()
    edg = topo["edg_vex"]
    assert np.all(edg[:, 0] < edg[:, 1])
    key = edg[:, 0].astype(np.int64) * (1 << 32) + edg[:, 1]
    assert np.all(np.diff(key) > 0)  # lexicographic, unique
    te = topo["tet_edg"]
    # faces: each tet has 4 faces = triples of its local edges
    faces = np.sort(np.stack([te[:, [0, 1, 3]], te[:, [0, 2, 4]], te[:, [1, 2, 5]], te[:, [3, 4, 5]]],
                             axis=1).reshape(-1, 3), axis=1)
    nF = len(np.unique(faces, axis=0))
    nV = len(np.unique(edg))
    assert nV - o.nE + nF - o.nT == 1


def test_tet_edges_and_signs_follow_vertex_order(oracle_mod_big, src_mod_big):
    """tet_edg(j) is the global edge of local edge j = (ind_tet(j,1),
    ind_tet(j,2)) and tet_drc(j) = +1 iff its first vertex id is smaller
    (src/FETD_FWD.f90:231-244)."""
    from temf3dt_amd import io as tfio
    m = tfio.read_mesh(src_mod_big["mesh"])
    t = oracle_mod_big.topology()
    tv = m["tet_vex"]
    for j, (p, q) in enumerate(IND_TET):
        va, vb = tv[:, p], tv[:, q]
        e = t["edg_vex"][t["tet_edg"][:, j] - 1]
        np.testing.assert_array_equal(e[:, 0], np.minimum(va, vb))
        np.testing.assert_array_equal(e[:, 1], np.maximum(va, vb))
        np.testing.assert_array_equal(t["tet_drc"][:, j], np.where(va < vb, 1, -1))


def _grads(x, y, z):
    a, b, c, d, vol = pyoracle.tet_cofactors(x, y, z)
    g = np.stack([b, c, d], axis=1) / (6.0 * vol)  # grad L_k
    return a, g, vol


def test_cofactors_unit_tet():
    x, y, z = UNIT
    a, b, c, d, vol = pyoracle.tet_cofactors(x, y, z)
    assert vol == pytest.approx(1.0 / 6.0, rel=1e-15)
    np.testing.assert_allclose(np.stack([b, c, d], 1), [[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], atol=1e-15)
    np.testing.assert_allclose(a, [1, 0, 0, 0], atol=1e-15)


def test_det_lu_against_numpy():
    rng = np.random.default_rng(3)
    for _ in range(50):
        m = rng.standard_normal((4, 4)) * 10 ** rng.uniform(-3, 3)
        assert pyoracle.lib().fo_det_lu(np.asfortranarray(m).ctypes.data_as(pyoracle._P), 4) == \
            pytest.approx(np.linalg.det(m), rel=1e-12)
    sing = np.array([[1.0, 2, 3, 1], [1, 2, 3, 1], [0, 1, 0, 1], [2, 0, 1, 1]])
    assert pyoracle.lib().fo_det_lu(np.asfortranarray(sing).ctypes.data_as(pyoracle._P), 4) == 0.0


def _whitney(x, y, z, pts):
    """Whitney functions N_i = l_i (L_a grad L_b - L_b grad L_a) at points."""
    a, g, vol = _grads(x, y, z)
    P = np.stack([x, y, z], 1)
    lam = (a[None, :] + pts @ (g * 6.0 * vol).T) / (6.0 * vol)
    N = []
    for (p, q) in IND_TET:
        li = np.linalg.norm(P[p] - P[q])
        N.append(li * (lam[:, p, None] * g[q] - lam[:, q, None] * g[p]))
    return np.stack(N, 1), g, vol


def test_mass_and_curl_matrices_brute_force():
    """f_kind_iso / e_kind_iso equal the Whitney integrals, checked by an exact
    quadrature (degree-2 four-point rule) on random tets."""
    rng = np.random.default_rng(11)
    aq = 0.5854101966249685
    bq = 0.1381966011250105
    bary = np.array([[aq, bq, bq, bq], [bq, aq, bq, bq], [bq, bq, aq, bq], [bq, bq, bq, aq]])
    for _ in range(5):
        P = rng.uniform(-3, 3, (4, 3))
        if np.linalg.det(np.c_[np.ones(4), P]) < 0:
            P[[2, 3]] = P[[3, 2]]
        x, y, z = P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()
        pts = bary @ P
        N, g, vol = _whitney(x, y, z, pts)
        Fq = vol / 4.0 * np.einsum("qid,qjd->ij", N, N)
        curl = np.stack([2 * np.linalg.norm(P[p] - P[q]) * np.cross(g[p], g[q]) for p, q in IND_TET])
        Kq = vol * curl @ curl.T
        F, K = pyoracle.tet_matrices(x, y, z, 1.0, 1.0)
        np.testing.assert_allclose(F, Fq, rtol=1e-12, atol=1e-14 * np.abs(Fq).max())
        np.testing.assert_allclose(K, Kq, rtol=1e-12, atol=1e-14 * np.abs(Kq).max())
        # K is singular on discrete gradients scaled by 1/l (edge DOFs are l * tangential E)
        phi = rng.standard_normal(4)
        lens = np.array([np.linalg.norm(P[p] - P[q]) for p, q in IND_TET])
        Gs = np.array([(phi[q] - phi[p]) for p, q in IND_TET]) / lens
        assert np.abs(K @ Gs).max() < 1e-10 * np.abs(K).max() * np.abs(Gs).max()
        assert np.linalg.eigvalsh(F).min() > 0


def test_source_waveform_is_derivative_of_erf_ramp():
    """erfstepon(1) = d/dt of 0.5 (1 + erf(5 (t - t0/2) / t0)) (type 1),
    type 3 is its negative (src/FETD_UTILITY.f90:1079-1095)."""
    t0 = 1e-7
    h = 1e-13
    for t in np.linspace(0, 2 * t0, 17):
        ramp = lambda s: 0.5 * (1 + erf(5 * (s - t0 / 2) / t0))
        num = (ramp(t + h) - ramp(t - h)) / (2 * h)
        assert pyoracle.source_didt(1, 1, t, t0, 1.0) == pytest.approx(num, rel=1e-5, abs=1e-3 / t0)
        assert pyoracle.source_didt(1, 3, t, t0, 2.0) == pytest.approx(-2 * pyoracle.source_didt(1, 1, t, t0, 1.0))


def _ward_hohmann_stepon(t, sig, a, I=1.0):
    """Central-loop dBz/dt of a step-on loop on a half-space (Ward & Hohmann
    1988 eq. 4.98 with the sign of the reference's z-down output column)."""
    mu0 = 4e-7 * math.pi
    th = np.sqrt(mu0 * sig / (4 * t))
    x = th * a
    return (I / (sig * a ** 3)) * (3 * erf(x) - 2 / np.sqrt(np.pi) * x * (3 + 2 * x * x) * np.exp(-x * x))


def test_oracle_against_analytic_halfspace(src_mod_big):
    """dBz/dt at the loop centre within 3% of the closed-form half-space
    response over 1e-5..1e-2 s (FE discretisation error, not rounding)."""
    g = np.load(src_mod_big["oracle_rows"])
    t, v = g["t"], g["fields"][:, 0, 5]
    for tq in [1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2]:
        k = int(np.argmin(abs(t - tq)))
        ratio = v[k] / _ward_hohmann_stepon(t[k], 1e-3, 500.0)
        assert 0.97 < ratio < 1.03, (tq, ratio)


def test_cntl_parser_and_required_keys(tmp_path, src_mod_big):
    from temf3dt_amd import io as tfio
    c = tfio.read_cntl(src_mod_big["cntl"])
    assert c["sp_division"] == 10 and c["sp_double"] == 12 and c["sp_dblesize"] == 2
    assert c["domains"][0] == [11.0, 1e-7, 1.0, 1.0]
    assert c["receivers"].tolist() == [[0.0, 0.0, 5.0]]
    bad = tmp_path / "bad.cntl"
    bad.write_text(open(src_mod_big["cntl"]).read().replace("sp_double: 12", "! no doubling"))
    with pytest.raises(ValueError, match="step_dble"):
        tfio.read_cntl(str(bad))


def test_oracle_missing_receiver_tet_message(tmp_path, src_mod_big):
    bad = tmp_path / "far.cntl"
    bad.write_text(open(src_mod_big["cntl"]).read().replace("0.0      0.0  5.0", "1.0e9 0.0 5.0"))
    with pytest.raises(RuntimeError, match="no rec tet"):
        pyoracle.Oracle(src_mod_big["mesh"], str(bad))


def test_anisotropic_matrices_brute_force():
    """f_kind_ani / e_kind_ani: the Whitney integrals with a tensor between
    the basis functions (mass) and between the curls (curl-curl), by exact
    quadrature on random tets with random symmetric tensors; with scalar
    tensors they reduce to the isotropic matrices."""
    rng = np.random.default_rng(12)
    aq, bq = 0.5854101966249685, 0.1381966011250105
    bary = np.array([[aq, bq, bq, bq], [bq, aq, bq, bq], [bq, bq, aq, bq], [bq, bq, bq, aq]])
    for _ in range(5):
        P = rng.uniform(-3, 3, (4, 3))
        if np.linalg.det(np.c_[np.ones(4), P]) < 0:
            P[[2, 3]] = P[[3, 2]]
        x, y, z = P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()
        B = rng.standard_normal((3, 3))
        Am = B @ B.T + 0.5 * np.eye(3)
        C = rng.standard_normal((3, 3))
        Ak = C + C.T
        N, g, vol = _whitney(x, y, z, bary @ P)
        Fq = vol / 4.0 * np.einsum("qid,de,qje->ij", N, Am, N)
        curl = np.stack([2 * np.linalg.norm(P[p] - P[q]) * np.cross(g[p], g[q]) for p, q in IND_TET])
        Kq = vol * curl @ Ak @ curl.T
        F, K = pyoracle.tet_matrices_ani(x, y, z, Am, Ak)
        np.testing.assert_allclose(F, Fq, rtol=1e-11, atol=1e-13 * np.abs(Fq).max())
        np.testing.assert_allclose(K, Kq, rtol=1e-11, atol=1e-13 * np.abs(Kq).max())
        Fi, Ki = pyoracle.tet_matrices(x, y, z, 2.5, 0.75)
        Fa, Ka = pyoracle.tet_matrices_ani(x, y, z, 2.5 * np.eye(3), 0.75 * np.eye(3))
        np.testing.assert_allclose(Fa, Fi, rtol=1e-14, atol=1e-15 * np.abs(Fi).max())
        np.testing.assert_allclose(Ka, Ki, rtol=1e-14, atol=1e-15 * np.abs(Ki).max())


def test_anisotropic_tensor_rotation():
    """model_ctl's tensor R diag(v) R^T, R = Rx Ry Rz from angles in degrees,
    with the reference's (2,2) entry of Rz taking the cosine of the second
    angle (src/FETD_FWD.f90:385-392); Python host code and oracle agree bit
    for bit."""
    from temf3dt_amd import io as tfio
    v = [1.0, 2.0, 3.0]
    # angles (0, 0, 90): Rz = [[0,-1,0],[1,cos 0 = 1,0],[0,0,1]]
    Rz = np.array([[0.0, -1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    np.testing.assert_allclose(pyoracle.aniso_tensor(v, [0, 0, 90]), Rz @ np.diag(v) @ Rz.T, atol=1e-15)
    # angles (30, 0, 0): a proper rotation about x
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    Rx = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    np.testing.assert_allclose(pyoracle.aniso_tensor(v, [30, 0, 0]), Rx @ np.diag(v) @ Rx.T, atol=1e-15)
    rng = np.random.default_rng(5)
    for _ in range(20):
        vv = 10.0 ** rng.uniform(-3, 1, 3)
        ang = rng.uniform(-180, 180, 3)
        assert np.array_equal(np.array(tfio.aniso_tensor(vv, ang, 2.0)), pyoracle.aniso_tensor(vv, ang, 2.0))


def test_anisotropic_cntl_parse(tmp_path):
    """The 'anisotropic:' section (three records per domain) gives the same
    per-tet tensors through the oracle and the Python host reader."""
    from temf3dt_amd import io as tfio
    from temf3dt_amd import workloads
    m = workloads.aniso_small_model()
    root = str(tmp_path / "ani")
    m.write(root)
    c = tfio.read_cntl(root + ".cntl")
    assert c["aniso"] and len(c["aniso"]) == len(m.aniso)
    mesh = tfio.read_mesh(root + ".mesh")
    s9, n9 = tfio.materials_tensor(mesh["tet_phy"], c["aniso"])
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    os9, on9 = o.tensors()
    o.close()
    assert np.array_equal(s9, os9)
    assert np.array_equal(n9, on9)


def test_impulse_waveform_is_derivative_of_its_gaussian():
    """Pulse type 2 (deltafunc(1), src/FETD_UTILITY.f90:1135-1165) is the
    time derivative of g(t) = exp(-((2e t - e t0) / t0)^2) (deltafunc(0)),
    by central differences and by the chain rule written out:
    g'(t) = -2 (2e t - e t0)(2e) / t0^2 g(t)."""
    e = math.e
    for t0 in (5e-9, 1e-7, 1e-6):
        h = 1e-7 * t0
        g = lambda s: math.exp(-(((2 * e) * s - e * t0) / t0) ** 2)  # noqa: E731
        for t in np.linspace(0.0, t0, 21):
            num = (g(t + h) - g(t - h)) / (2 * h)
            chain = -2.0 * ((2 * e) * t - e * t0) * (2 * e) / t0 ** 2 * g(t)
            got = pyoracle.source_didt(1, 2, t, t0, 1.0)
            scale = 4 * e / t0  # |g'| <= ~2.4 e / t0
            assert abs(got - chain) <= 1e-13 * scale, (t0, t)
            assert abs(got - num) <= 1e-7 * scale, (t0, t)
            # current scales the waveform
            assert pyoracle.source_didt(1, 2, t, t0, 3.0) == pytest.approx(3.0 * got, rel=1e-15, abs=0)
    # the pulse's integral over its support is g(t0) - g(0) = 0 (odd about t0/2)
    t0 = 1e-7
    ts = np.linspace(0.0, t0, 20001)
    vals = np.array([pyoracle.source_didt(1, 2, t, t0, 1.0) for t in ts])
    integ = np.sum(0.5 * (vals[1:] + vals[:-1]) * np.diff(ts))
    assert abs(integ) < 1e-9 * np.abs(vals).max() * t0


def test_field2_gate_rule_on_linear_rows():
    """The .field2 gate rule (src/FETD_FWD.f90:1624-1641): every gate is
    filled at the first row with t >= gate, by linear interpolation with the
    row before.  Rows linear in t are reproduced at the gate times; a gate
    already passed by the first row is skipped there (H12's `exit`) and
    filled at the second row, by extrapolation of that line."""
    rng = np.random.default_rng(11)
    t = np.cumsum(rng.uniform(0.5e-6, 2e-6, 40))
    t = np.concatenate([[0.0], t])
    alpha, beta = rng.standard_normal(9), rng.standard_normal(9)
    f = alpha[None, :] * t[:, None] + beta[None, :]
    gates = np.concatenate([[-1e-6, 0.0], np.sort(rng.uniform(t[1], t[-1], 12)), [t[7], 2 * t[-1]]])
    vals, row = pyoracle.gate_rows(t, f, gates)
    for g, tg in enumerate(gates):
        if tg > t[-1]:
            assert row[g] == -1
            continue
        want_row = max(1, int(np.argmax(t >= tg)))  # never the first row
        assert row[g] == want_row, (g, tg)
        np.testing.assert_allclose(vals[g], alpha * tg + beta, rtol=1e-12, atol=1e-12 * np.abs(beta).max())
    # a gate equal to a row time takes that row's values
    g7 = len(gates) - 2
    assert row[g7] == 7
    np.testing.assert_allclose(vals[g7], f[7], rtol=1e-13)


def test_field2_gate_rule_interpolates_between_bracketing_rows():
    """For gates inside (t[r-1], t[r]] the value is the chord between the
    bracketing rows (numpy's interp); several gates can fill on one row."""
    t = np.array([0.0, 1.0, 2.0, 4.0, 8.0, 16.0])
    f = np.stack([np.sin(t), t ** 2, np.exp(-t)], axis=1)
    gates = np.array([0.5, 1.5, 1.9, 3.0, 3.5, 5.0, 12.0])
    vals, row = pyoracle.gate_rows(t, f, gates)
    np.testing.assert_array_equal(row, [1, 2, 2, 3, 3, 4, 5])
    for q in range(3):
        np.testing.assert_allclose(vals[:, q], np.interp(gates, t, f[:, q]), rtol=1e-14, atol=1e-15)


def _stepoff_model(time_maximum):
    import dataclasses

    from temf3dt_amd import workloads
    m = workloads.small_model()
    return dataclasses.replace(m, time_maximum=time_maximum)


def _oracle_run(m, pulse_type, tmp_path, name, replay=None):
    import dataclasses
    st, _, t0, cur = m.source
    m = dataclasses.replace(m, source=(st, pulse_type, t0, cur))
    root = str(tmp_path / name)
    m.write(root)
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    o.assemble()
    if replay is not None:
        o.set_replay(replay)
    r = o.run()
    o.close()
    return r


@pytest.mark.parametrize("end", ["regular", "changed_check"])
def test_stepoff_second_pass_starts_from_first_pass_e1(end, tmp_path):
    """Two-pass step-off (src/FETD_FWD.f90:1212-1217,1295-1301,1453-1461):
    pass 1 negates the type-3 source, so it is exactly the step-on run
    (negation is exact in floating point); its final e1 starts pass 2 as
    both history states.  e1 is the last regular or kept solution: when the
    loop ends on a check that changed dt it is the regular step before that
    check, not the check's e_veri_1.  Pass 2's row 0 (E at the receivers
    from e1; dB/dt zeroed, H11) therefore equals a step-on run's E of the
    corresponding row, bit for bit, in both cases (one solver thread, so
    that PARDISO repeats its arithmetic exactly)."""
    pyoracle.set_threads(1)
    base = _stepoff_model(2e-4)
    on = _oracle_run(base, 1, tmp_path, "probe")
    chk_rows = [int(np.nonzero(on["step"] == s)[0][0]) for s in on["chk_step"]]
    if end == "regular":
        tmax = 0.5 * (on["t"][40] + on["t"][41])  # ends on a regular step
        last_src = None
    else:
        changed = [r for r, k in zip(chk_rows, on["chk_keep"]) if k == 0 and r > 1]
        assert changed, "the model must change dt at least once"
        r = changed[0]
        tmax = on["t"][r]  # the changed check's new time reaches time_maximum
        last_src = r
    m = _stepoff_model(tmax)
    on = _oracle_run(m, 1, tmp_path, "on")
    off = _oracle_run(m, 3, tmp_path, "off", replay=list(on["chk_keep"]) * 2)
    if end == "regular":
        assert on["step"][-1] % m.sp_double != 0
        src = len(on["step"]) - 1
    else:
        assert on["step"][-1] % m.sp_double == 0 and on["chk_keep"][-1] == 0
        src = len(on["step"]) - 2  # the regular step before the changed check
        assert src == last_src - 1
    assert off["step"][0] == 0
    np.testing.assert_array_equal(off["fields"][0, :, 0:3], on["fields"][src, :, 0:3])
    assert np.all(off["fields"][0, :, 3:6] == 0.0)
    if end == "changed_check":  # the check's own e_veri_1 differs from e1
        assert not np.array_equal(on["fields"][-1, :, 0:3], on["fields"][src, :, 0:3])


def test_stepoff_is_negated_stepon_after_the_ramp(tmp_path):
    """By linearity pass 2 = (free decay of e1) - (step-on response): with
    the same schedule, off + on is the decay of the pass-1 end state, which
    is small while t is small against time_maximum.  Tolerance: |dBz/dt of
    off + on| <= 2 % of |on| for rows past the ramp up to time_maximum / 20,
    and the sign flips (a dropped or doubled negation fails both)."""
    pyoracle.set_threads(4)
    m = _stepoff_model(2e-4)
    on = _oracle_run(m, 1, tmp_path, "on")
    off = _oracle_run(m, 3, tmp_path, "off", replay=list(on["chk_keep"]) * 2)
    np.testing.assert_array_equal(off["step"], on["step"])
    np.testing.assert_array_equal(off["t"], on["t"])
    a, b = off["fields"][:, 0, 5], on["fields"][:, 0, 5]
    sel = (on["step"] >= m.sp_division) & (on["t"] <= m.time_maximum / 20) & (np.abs(b) > 0)
    assert sel.sum() > 50
    rel = np.abs(a[sel] + b[sel]) / np.abs(b[sel])
    assert rel.max() < 0.02, rel.max()
    assert np.all(np.sign(a[sel]) == -np.sign(b[sel]))


def test_cpu_cost_sample(oracle_mod_big):
    """The CPU baseline's cost sample: phases 11/22/33 and the SpMV time on
    the shipped model, with the same factor size as the full run's phase 12."""
    pyoracle.set_threads(4)
    c = oracle_mod_big.cpu_costs(1e-8, nrhs_block=8)
    assert c["mtype"] == 2
    assert all(c[k] > 0 for k in ("phase11_s", "phase22_s", "phase33_s", "phase33_block_s", "rhs_s"))
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "src_mod_big", "oracle_rows.npz"))
    assert len(g["step"]) == 579
    r = oracle_mod_big.run(max_steps=1)
    assert int(c["factor_nnz"]) == int(r["factor_nnz"])


def test_oracle_matches_linux_build_at_air_1e12(tmp_path):
    """The reference's LINUX build wrote tests/golden/linux_mod_big's .field
    for the mod_big mesh with air at 1e-12 S/m (the conditioning of the
    3-D configurations) and 4 printed digits.  It is a different revision
    of adap_euler: after a dt change its time column advances by another rule
    (0.2200e-06 against src's 0.21667e-06 at step 64), so the two schedules
    part there.  Up to that first change (64 rows: the 30-step source ramp,
    regular steps and a kept check at step 32) the oracle reproduces its
    steps, times and dBz/dt to the printed digits (relative 5e-4)."""
    import gzip
    import shutil
    d = os.path.join(os.path.dirname(__file__), "golden", "linux_mod_big")
    mesh = str(tmp_path / "m.mesh")
    with gzip.open(os.path.join(d, "mod_big_sig_0.001.mesh.gz"), "rb") as f, open(mesh, "wb") as g:
        shutil.copyfileobj(f, g)
    ref = np.loadtxt(os.path.join(d, "mod_big_sig_0.001.field"))
    t = ref[:, 1]
    first = next(i for i in range(2, len(t)) if (t[i] - t[i - 1]) / (t[i - 1] - t[i - 2]) > 1.5)
    assert first == 64
    pyoracle.set_threads(os.cpu_count() or 8)
    o = pyoracle.Oracle(mesh, os.path.join(d, "mod_big_sig_0.001.cntl"))
    o.assemble()
    r = o.run(max_steps=first - 1)
    o.close()
    assert len(r["step"]) == first
    np.testing.assert_array_equal(r["step"], ref[:first, 0])
    np.testing.assert_allclose(r["t"][1:], t[1:first], rtol=5e-4)
    a, b = r["fields"][:, 0, 5], ref[:first, 2]
    sel = b != 0
    assert np.all(a[~sel] == 0.0)
    np.testing.assert_allclose(a[sel], b[sel], rtol=5e-4)
