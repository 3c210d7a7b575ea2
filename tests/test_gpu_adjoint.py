"""Adjoint gradient of a receiver-row functional with respect to per-tet
log-conductivity (tf_adjoint): against central finite differences of the GPU
forward and of the oracle (the reference algorithm on the CPU), on a fixed
(replayed) schedule; independence of the checkpoint interval; the
torch.autograd.Function wrapper."""
import os

import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu

H = 1e-4  # central difference step in log(sigma)


@pytest.fixture(scope="module")
def adj_setup():
    from temf3dt_amd import workloads
    m = workloads.small_model()
    fe = gpu_from_model(m)
    sig, imu = m.sigma_inv_mu()
    base = fe.run()
    keep = fe.checks()["keep"]
    fe.set_replay(keep)
    f = base["fields"]
    rng = np.random.default_rng(4)
    # weights on dB/dt and B at every receiver and on E at the receiver in
    # the earth, scaled to each component's size so each one matters.  (E in
    # the 1e-8 S/m air carries a gradient part that only the tiny air
    # conductivity pins; its value there moves with rounding in the
    # factorisation, which finite differences would pick up as noise.)
    scale = np.sqrt((f ** 2).mean(axis=0)) + 1e-300
    W = rng.standard_normal(f.shape) / scale / f.shape[0]
    in_air = m.receivers[:, 2] < 0
    W[:, in_air, 0:3] = 0.0
    cent = m.xyz[m.tet_vex - 1].mean(axis=1)
    near = np.flatnonzero((cent[:, 2] > 0) & (np.abs(cent[:, 0]) < 150) & (np.abs(cent[:, 1]) < 150)
                          & (cent[:, 2] < 150))
    cells = np.sort(rng.choice(near, 8, replace=False))
    yield dict(m=m, fe=fe, sig=sig, imu=imu, keep=keep, W=W, cells=cells, nrows=f.shape[0])
    fe.close()


def _J(fe, sig, imu, W):
    fe.set_materials(sig, imu)
    fe.assemble()
    f = fe.run()["fields"]
    assert f.shape == W.shape
    return float((W * f).sum())


def _adjoint(s, every):
    fe = s["fe"]
    fe.set_materials(s["sig"], s["imu"])
    fe.assemble()
    fe.set_checkpointing(every)
    rows = fe.run()
    g = fe.adjoint(s["W"])
    fe.set_checkpointing(0)
    return rows, g


def test_adjoint_matches_gpu_finite_differences(adj_setup):
    s = adj_setup
    print(f"[gpu] adjoint: {s['m'].ntet} tets, {s['nrows']} rows, cells {list(s['cells'])}", flush=True)
    _, g = _adjoint(s, 50)
    assert np.all(np.isfinite(g))
    for t in s["cells"]:
        sp, sm = s["sig"].copy(), s["sig"].copy()
        sp[t] *= np.exp(H)
        sm[t] *= np.exp(-H)
        fd = (_J(s["fe"], sp, s["imu"], s["W"]) - _J(s["fe"], sm, s["imu"], s["W"])) / (2 * H)
        rel = abs(g[t] - fd) / max(abs(fd), 1e-300)
        print(f"[gpu] cell {t}: adjoint {g[t]:.6e} fd {fd:.6e} rel {rel:.2e}", flush=True)
        assert rel < 1e-4, (t, g[t], fd)


def test_adjoint_independent_of_checkpoint_interval(adj_setup):
    s = adj_setup
    _, g50 = _adjoint(s, 50)
    _, g7 = _adjoint(s, 7)
    _, gall = _adjoint(s, 100000)
    scale = np.abs(g50).max()
    print(f"[gpu] ckpt 7 vs 50: {np.abs(g7 - g50).max() / scale:.2e}, one segment vs 50: "
          f"{np.abs(gall - g50).max() / scale:.2e}", flush=True)
    assert np.abs(g7 - g50).max() <= 1e-10 * scale
    assert np.abs(gall - g50).max() <= 1e-10 * scale


def test_adjoint_factoring_ahead_is_bit_identical(adj_setup):
    """The sweep's next dt factored on the second stream while the current
    one is in use gives the gradient of factoring at the point of use, bit
    for bit (the same factors, the same order of operations)."""
    s = adj_setup
    fe = s["fe"]
    _, g1 = _adjoint(s, 7)
    fe.set_tuning("overlap", 0)
    try:
        _, g0 = _adjoint(s, 7)
    finally:
        fe.set_tuning("overlap", 1)
    np.testing.assert_array_equal(g1, g0)


def test_adjoint_with_kept_factors_is_bit_identical(adj_setup):
    """The taped run keeps the factors it would overwrite, and the sweep
    swaps them into its slots instead of factoring P(dt) again: the same
    gradient, bit for bit, with fewer factorisations once the run used more
    dt than the two slots hold."""
    s = adj_setup
    fe = s["fe"]
    rows, g1 = _adjoint(s, 7)
    n1 = fe.timers()["nfactor"]
    counts = {}
    for cap in (1, 0):  # at most one kept (the sweep mixes kept and refactored dt), none kept
        fe.set_tuning("keep_factors", cap)
        try:
            _, g0 = _adjoint(s, 7)
            counts[cap] = fe.timers()["nfactor"]
        finally:
            fe.set_tuning("keep_factors", 8)
        np.testing.assert_array_equal(g1, g0)
    ndt = len(np.unique(rows["dt"]))
    print(f"[gpu] kept factors: {ndt} dt in the run, factorisations {n1:.0f} (up to 8 kept), "
          f"{counts[1]:.0f} (1 kept), {counts[0]:.0f} (none kept)", flush=True)
    assert n1 <= counts[1] <= counts[0]
    if ndt > 2:
        assert n1 < counts[0]


def test_adjoint_with_kept_states_is_bit_identical(adj_setup):
    """The taped run keeps its nodes' solutions and the sweep reads them
    instead of recomputing each checkpoint segment: the same gradient, bit
    for bit, as the sweep that recomputes every state."""
    s = adj_setup
    fe = s["fe"]
    _, g1 = _adjoint(s, 7)
    fe.set_tuning("keep_states", 0)
    try:
        _, g0 = _adjoint(s, 7)
    finally:
        fe.set_tuning("keep_states", 1)
    np.testing.assert_array_equal(g1, g0)


def test_adjoint_matches_oracle_finite_differences(adj_setup, tmp_path):
    """Central differences through the oracle (the reference algorithm with
    MKL PARDISO on the CPU) on the same replayed schedule."""
    from oracle import pyoracle
    s = adj_setup
    _, g = _adjoint(s, 50)
    root = str(tmp_path / "small")
    s["m"].write(root)
    pyoracle.set_threads(min(16, os.cpu_count() or 1))

    def j_oracle(sig):
        o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
        o.set_sigma(sig)
        o.assemble()
        o.set_replay(s["keep"])
        f = o.run()["fields"]
        o.close()
        assert f.shape == s["W"].shape
        return float((s["W"] * f).sum())

    for t in s["cells"][:2]:
        sp, sm = s["sig"].copy(), s["sig"].copy()
        sp[t] *= np.exp(H)
        sm[t] *= np.exp(-H)
        fd = (j_oracle(sp) - j_oracle(sm)) / (2 * H)
        rel = abs(g[t] - fd) / max(abs(fd), 1e-300)
        print(f"[gpu] cell {t}: adjoint {g[t]:.6e} oracle fd {fd:.6e} rel {rel:.2e}", flush=True)
        assert rel < 1e-3, (t, g[t], fd)


def test_torch_autograd_wrapper(adj_setup):
    import torch
    from temf3dt_amd import torch_fetd
    s = adj_setup
    fe = s["fe"]
    ls = torch.tensor(np.log(s["sig"]), dtype=torch.float64, requires_grad=True)
    fields = torch_fetd.fetd_fields(ls, fe, s["imu"], ckpt_every=50, replay=s["keep"])
    loss = (torch.from_numpy(s["W"]) * fields).sum()
    loss.backward()
    _, g = _adjoint(s, 50)
    err = np.abs(ls.grad.numpy() - g).max() / np.abs(g).max()
    print(f"[gpu] torch wrapper vs tf_adjoint: {err:.2e}", flush=True)
    assert err < 1e-12
    fe.set_checkpointing(0)
