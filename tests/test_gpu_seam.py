"""The PARDISO seam of the reference, driven as the reference drives it.

adap_euler (src/FETD_FWD.f90:1229-1464) keeps its matrices, right-hand
sides, SpMV and controller on the host and calls the sparse solver through
pardiso_direct_1 (phase 12, :1307,1361), pardiso_direct_2 (phase 33,
:1334,1371,1431) and pardiso_direct_3 (phase -1, :1420,1446).  This test
restates that caller in Python over the CPU-assembled matrices of the
shipped src/mod_big_sig_0.001 model (the oracle's mesh_process output) and
sends only those three calls to the library: tf_factor_csr, tf_solve and
tf_release.  The rows must reproduce the committed oracle rows.

It also checks the seam's failure path: the reference retries a failed SPD
factorisation as symmetric indefinite (src/FETD_UTILITY.f90:919-935); here
a P with negative pivots is reported as TF_ERR_NUMERIC, or factored when the
caller accepts indefinite factors.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from temf3dt_amd import fetd, io as tfio

pytestmark = pytest.mark.gpu


def _full(ia, ja, va):
    n = len(ia) - 1
    U = sp.csr_matrix((va, ja - 1, ia - 1), shape=(n, n))
    return (U + U.T - sp.diags(U.diagonal())).tocsr()


def _seam_context(src_mod_big):
    m = tfio.read_mesh(src_mod_big["mesh"])
    fe = fetd.Fetd(0)
    fe.set_mesh(m["xyz"], m["tet_vex"])  # the pattern the values are checked against
    return fe


def test_reference_loop_through_pardiso_seam(src_mod_big, oracle_mod_big):
    from oracle import pyoracle
    o = oracle_mod_big
    csr, top, rec, src = o.csr(), o.topology(), o.receivers(), o.source()
    c = tfio.read_cntl(src_mod_big["cntl"])
    sp_division, sp_double, dbl = c["sp_division"], c["sp_double"], c["sp_dblesize"]
    t0, cur, tmax = c["pulse_width"], c["current"], c["time_maximum"]
    ia, ja, B, C = csr["ia"], csr["ja"], csr["mass"], csr["curl"]
    n = len(ia) - 1
    Bf = _full(ia, ja, B)
    edg_len = top["edg_len"]
    se, sd = src["src_edg"] - 1, src["src_drc"]
    te = top["tet_edg"][rec["rec_tet"] - 1] - 1      # [R][6]
    tdrc = top["tet_drc"][rec["rec_tet"] - 1]         # [R][6]

    fe = _seam_context(src_mod_big)
    rows_t, rows_step, rows_f = [], [], []

    def output(step, t, dt, e):  # solution_output, receiver part (:1532-1541,1602-1620)
        ev = e[te] * tdrc
        E = np.einsum("rjk,rj->rk", rec["n_arr"], ev)
        D = np.einsum("rjk,rj->rk", rec["d_arr"], ev)
        if step == 0:
            D[:] = 0.0
        Bc = D * dt if not rows_f else rows_f[-1][:, 6:9] + D * dt
        rows_step.append(step)
        rows_t.append(t)
        rows_f.append(np.concatenate([E, D, Bc], axis=1))

    def rhs(step, t, e2, e3, dt):  # sys_right + SpMV (:1466-1511, :1427-1428)
        s = np.zeros(n)
        if step < sp_division:
            s[se] = pyoracle.source_didt(1, 1, t, t0, cur) * sd * edg_len[se]
        return Bf @ (4.0 * e2 - e3) - 2.0 * dt * s

    def factor(slot, dt):  # mkl_add_3 + pardiso_direct_1
        fe.factor_csr(slot, ia, ja, 2.0 * dt * C + 3.0 * B)

    def solve(slot, r):  # pardiso_direct_2
        return fe.solve(slot, r)

    step, not_changed, keep_log = 0, 0, []
    time_cur, step_cur = 0.0, t0 / sp_division
    e2, e3 = np.zeros(n), np.zeros(n)
    alle = []
    ind = 0
    nfact = 0
    while time_cur < tmax:
        step += 1
        time_cur = time_cur + step_cur
        if step == 1:
            output(0, time_cur - step_cur, step_cur, e2)
            factor(ind, step_cur)
            nfact += 1
        if step % sp_double == 0:
            te2, te3, ttime, tstep = e2, e3, time_cur, step_cur
            for i in range(1, dbl + 1):
                te1 = solve(ind, rhs(step, ttime, te2, te3, tstep))
                if i == dbl:
                    break
                if i == 1:
                    erec = te1
                te3, te2 = te2, te1
                ttime = ttime + tstep
            ev1 = te1
            tstep = step_cur * dbl
            ttime = time_cur - step_cur + tstep
            other = 1 - ind
            factor(other, tstep)
            nfact += 1
            ev2 = solve(other, rhs(step, ttime, alle[-1], alle[-1 - dbl], tstep))
            err = np.linalg.norm(ev2 - ev1) / np.linalg.norm(ev1)
            keep = err > 1e-4 and not_changed < 2
            keep_log.append(int(keep))
            if keep:
                fe.release(other)  # pardiso_direct_3 of the trial factor
                e1 = erec
                output(step, time_cur, step_cur, e1)
                e3, e2 = e2, e1
                not_changed += 1
            else:
                fe.release(ind)
                ind = other
                e3, e2 = alle[-1], ev1
                time_cur, step_cur = ttime, tstep
                output(step, time_cur, step_cur, e2)
                not_changed = 0
            alle = []
            continue
        e1 = solve(ind, rhs(step, time_cur, e2, e3, step_cur))
        output(step, time_cur, step_cur, e1)
        alle.append(e1)
        e3, e2 = e2, e1
    fe.release(ind)
    piv = fe.factor_pivots()
    fe.close()

    g = np.load(src_mod_big["oracle_rows"])
    print(f"[gpu] seam loop: {len(rows_t)} rows, {nfact} factorisations, last pivots {piv}", flush=True)
    np.testing.assert_array_equal(np.array(rows_step), g["step"])
    np.testing.assert_array_equal(np.array(rows_t), g["t"])
    np.testing.assert_array_equal(keep_log, g["chk_keep"])
    assert nfact == 49  # one at step 1 and one at every check, as the reference does
    assert piv["negative"] == 0
    a = np.array(rows_f)[:, 0, 5]
    b = g["fields"][:, 0, 5]
    sel = np.abs(b) > 0
    rel = np.abs(a[sel] - b[sel]) / np.abs(b[sel])
    print(f"[gpu] seam loop dBz/dt max rel diff vs oracle {rel.max():.3e}", flush=True)
    assert rel.max() < 1e-6


def test_seam_solution_matches_pardiso(src_mod_big, oracle_mod_big):
    """One tf_factor_csr / tf_solve against MKL PARDISO on the same P and b."""
    from oracle import pyoracle
    if not pyoracle.pardiso_available():
        pytest.skip("MKL PARDISO is not available on this host")
    csr = oracle_mod_big.csr()
    va = 2.0 * 1e-8 * csr["curl"] + 3.0 * csr["mass"]
    fe = _seam_context(src_mod_big)
    fe.factor_csr(0, csr["ia"], csr["ja"], va)
    b = np.random.default_rng(5).standard_normal(len(csr["ia"]) - 1)
    x = fe.solve(0, b)
    fe.close()
    xp = pyoracle.pardiso_solve(csr["ia"], csr["ja"], va, b)
    rel = np.linalg.norm(x - xp) / np.linalg.norm(xp)
    print(f"[gpu] tf_solve vs PARDISO rel diff {rel:.3e}", flush=True)
    assert rel < 1e-8


def test_seam_indefinite_matrix_reports_numeric_error(src_mod_big, oracle_mod_big):
    """P with one sign-flipped block: the edges S = {1..n/10} are decoupled
    from the rest and their block negated, P' = blockdiag(-P_SS, P_NN).  By
    Sylvester's law of inertia every symmetric LDL^T of P' has exactly |S|
    negative pivots (each block is definite, so no pivot is small).
    tf_factor_csr returns TF_ERR_NUMERIC naming them; with tuning key
    "indefinite" the unpivoted LDL^T is accepted and solves the system."""
    csr = oracle_mod_big.csr()
    ia, ja = csr["ia"], csr["ja"]
    n = len(ia) - 1
    va = 2.0 * 1e-6 * csr["curl"] + 3.0 * csr["mass"]
    rows = np.repeat(np.arange(n), np.diff(ia))
    cols = ja - 1
    ns = n // 10
    in_r, in_c = rows < ns, cols < ns
    va_bad = va.copy()
    va_bad[in_r & in_c] = -va[in_r & in_c]
    va_bad[in_r != in_c] = 0.0
    fe = _seam_context(src_mod_big)
    fe.factor_csr(0, ia, ja, va)
    assert fe.factor_pivots()["negative"] == 0
    with pytest.raises(fetd.FetdError) as ex:
        fe.factor_csr(1, ia, ja, va_bad)
    assert ex.value.code == fetd.TF_ERR_NUMERIC
    assert "negative pivots" in str(ex.value)
    neg = fe.factor_pivots()["negative"]
    print(f"[gpu] flipped P: {neg} negative pivots reported for |S| = {ns}", flush=True)
    assert neg == ns
    fe.set_tuning("indefinite", 1)
    fe.factor_csr(1, ia, ja, va_bad)
    assert fe.factor_pivots()["negative"] == neg
    b = np.random.default_rng(3).standard_normal(n)
    x = fe.solve(1, b)
    fe.close()
    Px = _full(ia, ja, va_bad) @ x
    berr = np.abs(Px - b).max() / (np.abs(va_bad).max() * np.abs(x).max() + np.abs(b).max())
    print(f"[gpu] indefinite factor backward error {berr:.3e}", flush=True)
    assert berr < 1e-10
