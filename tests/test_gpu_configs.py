"""GPU parity at the configurations' sizes, with air at 1e-12 S/m.

The oracle rows (tests/golden/*/oracle_rows.npz, written by
tests/golden/make_config_rows.py with the CPU restatement and MKL PARDISO)
cover the first 110 steps of each model: the source ramp, regular steps and
the first doubling check at step 100 with its refactorisation.

  plain 3-D        924,768 tets, 1.1M edges, 41 receivers, one source
  batched 3-D      1,346,700 tets, 64 transmitters in one block solve;
                   transmitters 0 and 45 against single-source oracle runs on
                   the same (replayed) schedule
  batched 3-D,     the same transmitters and earth on the mesh bisected to
  refined          1.25 m cells at every wire (1,690,764 tets), likewise
  uneven models    the 48 seeded models in one context; the smallest, the
                   median and the largest against their oracle runs
  adjoint          the plain 3-D mesh, 160 steps with checkpoints every 50
                   (4 segments): the misfit gradient against central finite
                   differences of the GPU forward

dB/dt must agree within 1e-6 relative wherever it is
above 1e-6 of its largest value, with identical steps, times and check
decisions, and every factorisation must be positive definite (no negative
pivot; a negative one would fail the run with TF_ERR_NUMERIC).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, gpu_from_model

pytestmark = pytest.mark.gpu

STEPS = 110


def _golden(sub):
    return np.load(os.path.join(GOLDEN, sub, "oracle_rows.npz"))


def _fingerprint(m):
    return np.array([m.ntet, len(m.xyz), float(np.sum(m.xyz))])


def _compare_dbdt(a, b, what):
    """a, b: [..., 9] rows; dB/dt (columns 3-5) within 1e-6 relative where
    above 1e-6 of the largest value, and within 1e-6 of it everywhere."""
    a, b = a[..., 3:6], b[..., 3:6]
    scale = np.abs(b).max()
    big = np.abs(b) > 1e-6 * scale
    rel = np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))
    absd = np.abs(a - b).max() / scale
    print(f"[gpu] {what}: max rel dB/dt diff {rel:.3e} over {big.sum()} values, max abs / scale {absd:.3e}",
          flush=True)
    assert rel < 1e-6
    assert absd < 1e-6


def test_cfg_plain3d_matches_oracle():
    from temf3dt_amd import workloads
    m = workloads.plain_3d_model()
    g = _golden("plain3d")
    np.testing.assert_allclose(_fingerprint(m), g["fp"], rtol=1e-12)
    print(f"[gpu] plain 3-D: {m.ntet} tets, {STEPS} steps", flush=True)
    fe = gpu_from_model(m)
    assert fe.nE == int(g["nE"])
    rows = fe.run(max_steps=STEPS)
    ck = fe.checks()
    piv = fe.factor_pivots()
    st = fe.factor_stats()
    fe.close()
    print(f"[gpu] plain 3-D: nnzL {st['nnzL']:.3e} (PARDISO {int(g['factor_nnz']):.3e}), pivots {piv}", flush=True)
    assert piv["negative"] == 0 and piv["min_abs"] > 0
    np.testing.assert_array_equal(rows["step"], g["step"])
    np.testing.assert_array_equal(rows["t"], g["t"])
    np.testing.assert_array_equal(ck["keep"], g["chk_keep"])
    np.testing.assert_allclose(ck["err"], g["chk_err"], rtol=1e-3)
    _compare_dbdt(rows["fields"], g["fields"], "plain 3-D, 41 receivers")  # [row][receiver][9]


@pytest.mark.parametrize("name", ["batched3d", "batched3d_refined"])
def test_cfg_batched3d_transmitters_match_oracle(name):
    from temf3dt_amd import workloads
    print(f"[gpu] {name}: building the mesh", flush=True)
    m = {"batched3d": workloads.batched_3d_model, "batched3d_refined": workloads.batched_3d_refined_model}[name]()
    g = _golden(name)
    np.testing.assert_allclose(_fingerprint(m), g["fp"], rtol=1e-12)
    print(f"[gpu] {name}: {m.ntet} tets, {len(m.loops)} transmitters, {STEPS} steps", flush=True)
    fe = gpu_from_model(m)
    assert fe.nE == int(g["nE"])
    fe.set_replay(list(g["chk_keep"]))
    rows = fe.run(max_steps=STEPS)
    piv = fe.factor_pivots()
    fe.close()
    assert piv["negative"] == 0
    np.testing.assert_array_equal(rows["step"], g["step"])
    np.testing.assert_array_equal(rows["t"], g["t"])
    for i, tx in enumerate(g["tx"]):
        ref = g[f"fields_{i}"]  # [row][receiver][9], transmitter tx alone
        _compare_dbdt(rows["fields"][:, tx], ref, f"{name} transmitter {tx}, all 64 receivers")
        _compare_dbdt(rows["fields"][:, tx, tx], ref[:, tx], f"{name} transmitter {tx}, coincident receiver")


def test_cfg_uneven48_models_match_oracle():
    from temf3dt_amd import fetd, workloads
    ms = workloads.uneven_models(48)
    g = _golden("uneven48")
    np.testing.assert_array_equal([mm.ntet for mm in ms], g["sizes"])
    u = workloads.union(ms)
    print(f"[gpu] 48 uneven models: {int(u['ntet'].sum())} tets in one context", flush=True)
    fe = fetd.Fetd(0)
    fe.set_mesh(u["xyz"], u["tet_vex"])
    fe.set_models(u["ntet"], u["nrec"])
    fe.set_materials(u["sigma"], u["inv_mu"])
    m0 = ms[0]
    for i, lp in enumerate(u["loops"]):
        fe.set_source(lp, *m0.source, isrc=i)
    fe.set_receivers(u["receivers"], max(mm.rec_charac for mm in ms))
    fe.set_schedule(m0.sp_division, m0.sp_double, m0.sp_dblesize, m0.time_maximum)
    fe.assemble()
    fe.run_models(max_steps=STEPS, graphs=True)
    piv = fe.factor_pivots()
    got = {int(k): fe.model_rows(int(k)) for k in g["pick"]}
    fe.close()
    assert piv["negative"] == 0
    for i, k in enumerate(g["pick"]):
        k = int(k)
        np.testing.assert_allclose(_fingerprint(ms[k]), g[f"fp_{i}"], rtol=1e-12)
        r = got[k]
        np.testing.assert_array_equal(r["step"], g[f"step_{i}"])
        np.testing.assert_array_equal(r["t"], g[f"t_{i}"])
        _compare_dbdt(r["fields"], g[f"fields_{i}"], f"uneven model {k} ({ms[k].ntet} tets)")


def _gates_and_weights(t, gates):
    """Rows filling each gate by the .field2 rule (first row r >= 1 with
    t[r] >= gate) and the interpolation weights on rows r-1 and r."""
    out = []
    for tg in gates:
        r = int(np.argmax(t >= tg))
        r = max(r, 1)
        w1 = (t[r] - tg) / (t[r] - t[r - 1])
        out.append((r, w1, 1.0 - w1))
    return out


def test_cfg_adjoint_gradient_on_plain3d_mesh():
    """Misfit J = 1/2 sum over gates and receivers of ((d - d_obs) / s)^2 of
    dB/dt at 41 receivers x 40 gates (log-spaced over the bounded run),
    d_obs from the 3-block true model with 2 % Gaussian noise (seed 4),
    s = 2 % |d_obs| + 1e-3 max|d_obs|; the model is the 0.01 S/m half-space.
    The adjoint gradient with respect to log sigma of 4 cells (the largest
    |gradient| among 16 seeded cells near the receiver line) agrees with
    central differences (h = 1e-4) of the GPU forward on the frozen schedule
    within 1e-3."""
    import dataclasses

    from temf3dt_amd import workloads
    steps, ckpt = 160, 50
    base = workloads.plain_3d_model()
    true_sig = workloads.tet_sigma(base, 0.01, workloads.adjoint_blocks(4))
    sig0 = workloads.tet_sigma(base, 0.01)
    _, imu = base.sigma_inv_mu()
    m = dataclasses.replace(base)
    print(f"[gpu] adjoint: {m.ntet} tets, {steps} steps, checkpoints every {ckpt}", flush=True)
    fe = gpu_from_model(m)
    fe.set_materials(true_sig, imu)
    fe.assemble()
    obs_rows = fe.run(max_steps=steps)
    keep = fe.checks()["keep"]
    t = obs_rows["t"]
    gates = np.logspace(np.log10(2e-6), np.log10(0.98 * t[-1]), 40)
    gw = _gates_and_weights(t, gates)

    def gate_vals(rows):
        f = rows["fields"][:, :, 5]  # dBz/dt [row][receiver] (one source)
        return np.array([w1 * f[r - 1] + w2 * f[r] for r, w1, w2 in gw])

    d_true = gate_vals(obs_rows)
    rng = np.random.default_rng(4)
    d_obs = d_true * (1.0 + 0.02 * rng.standard_normal(d_true.shape))
    sd = 0.02 * np.abs(d_obs) + 1e-3 * np.abs(d_obs).max()

    fe.set_replay(keep)

    def misfit(sig):
        fe.set_materials(sig, imu)
        fe.assemble()
        rows = fe.run(max_steps=steps)
        np.testing.assert_array_equal(rows["t"], t)
        return 0.5 * np.sum(((gate_vals(rows) - d_obs) / sd) ** 2), rows

    fe.set_materials(sig0, imu)
    fe.assemble()
    fe.set_checkpointing(ckpt)
    rows = fe.run(max_steps=steps)
    resid = (gate_vals(rows) - d_obs) / sd ** 2  # dJ/d(gate value)
    w = np.zeros_like(rows["fields"])  # [row][rec][9]
    for (r, w1, w2), res in zip(gw, resid):
        w[r - 1, :, 5] += w1 * res
        w[r, :, 5] += w2 * res
    grad = fe.adjoint(w)
    fe.set_checkpointing(0)
    assert np.all(np.isfinite(grad))
    # candidate cells: earth tets near the receiver line, seeded
    c = m.xyz[m.tet_vex - 1].mean(axis=1)
    near = np.nonzero((np.abs(c[:, 0]) < 380) & (np.abs(c[:, 1]) < 60) & (c[:, 2] > 0) & (c[:, 2] < 80))[0]
    cand = np.random.default_rng(4).choice(near, 16, replace=False)
    cells = cand[np.argsort(-np.abs(grad[cand]))[:4]]
    h = 1e-4
    worst = 0.0
    for cell in cells:
        sp, sm = sig0.copy(), sig0.copy()
        sp[cell] *= np.exp(h)
        sm[cell] *= np.exp(-h)
        jp, _ = misfit(sp)
        jm, _ = misfit(sm)
        fd = (jp - jm) / (2 * h)
        rel = abs(fd - grad[cell]) / abs(fd)
        worst = max(worst, rel)
        print(f"[gpu] adjoint cell {cell}: adjoint {grad[cell]:.6e} central FD {fd:.6e} rel {rel:.2e}", flush=True)
    fe.close()
    assert worst < 1e-3
