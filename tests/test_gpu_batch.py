"""Block solve (K right-hand sides on one factorisation) and batched
multi-transmitter time stepping on the GPU, against the single-RHS path and
the oracle."""
import os

import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu


def _upper_matvec(ia, ja, va, x):
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    y = np.zeros_like(x)
    np.add.at(y, rows, va[:, None] * x[ja]) if x.ndim == 2 else np.add.at(y, rows, va * x[ja])
    off = rows != ja
    if x.ndim == 2:
        np.add.at(y, ja[off], va[off, None] * x[rows[off]])
    else:
        np.add.at(y, ja[off], va[off] * x[rows[off]])
    return y


@pytest.mark.parametrize("nrhs", [5, 64])
def test_block_solve_matches_single_rhs(gpu_mod_big_batch, nrhs):
    fe, oc = gpu_mod_big_batch
    rng = np.random.default_rng(nrhs)
    B = rng.standard_normal((nrhs, fe.nE))
    print(f"[gpu] block solve nrhs={nrhs}", flush=True)
    X = fe.solve(0, B)
    for q in [0, nrhs - 1]:
        x1 = fe.solve(0, B[q])
        rel = np.linalg.norm(X[q] - x1) / np.linalg.norm(x1)
        assert rel < 1e-10, (q, rel)
    va = 2.0 * 1e-8 * oc["curl"] + 3.0 * oc["mass"]
    ia, ja = oc["ia"] - 1, oc["ja"] - 1
    R = _upper_matvec(ia, ja, va, X.T) - B.T
    berr = np.abs(R).max() / (np.abs(va).max() * np.abs(X).max() + np.abs(B).max())
    print(f"[gpu] block backward error {berr:.3e}", flush=True)
    assert berr < 1e-12


def test_block_solve_row_major_l21_matches_single_rhs():
    """A multi-source context keeps a row-major copy of every front's L21 for
    the block solve's backward product (N form over struct rows); its block
    solves agree with the single-RHS path, which reads L21 itself."""
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    fe = gpu_from_model(m)  # four sources
    fe.factor(0, 1e-8)
    rng = np.random.default_rng(11)
    for nrhs in (4, 64):
        B = rng.standard_normal((nrhs, fe.nE))
        print(f"[gpu] row-major L21 block solve nrhs={nrhs}", flush=True)
        X = fe.solve(0, B)
        for q in (0, nrhs - 1):
            x1 = fe.solve(0, B[q])
            rel = np.linalg.norm(X[q] - x1) / np.linalg.norm(x1)
            assert rel < 1e-10, (nrhs, q, rel)
    fe.close()


@pytest.mark.parametrize("sources", [[0, 1, 2, 3], [1]])
def test_partitioned_inverse_matches_l21(sources):
    """The lower levels' partitioned inverse (L21 replaced by -L21 L11^-1 at
    factorisation; one tile product per front and direction) against the
    plain L21 factor: block solves with and without the row-major copy (four
    sources / one source) and the single right-hand-side solve agree to
    rounding and keep an FP64 backward error."""
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    fe = gpu_from_model(m, loops=sources)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((64, fe.nE))
    out = {}
    try:
        for pv in (0, 10):
            print(f"[gpu] partitioned inverse: sources {sources}, pinv={pv}", flush=True)
            fe.set_tuning("pinv", pv)
            fe.factor(0, 1e-8)
            out[pv] = (fe.solve(0, B), fe.solve(0, B[:20]), fe.solve(0, B[3]))
    finally:
        fe.set_tuning("pinv", 10)
    for i, name in enumerate(("64 rhs", "20 rhs", "1 rhs")):
        a, b = out[0][i], out[10][i]
        rel = np.abs(b - a).max() / np.abs(a).max()
        print(f"[gpu] pinv vs L21, {name}: {rel:.3e}", flush=True)
        assert rel < 1e-10
    fe.close()


@pytest.fixture(scope="module")
def gpu_mod_big_batch(src_mod_big, oracle_mod_big):
    from temf3dt_amd import fetd
    fe, _, _ = fetd.setup_from_files(src_mod_big["mesh"], src_mod_big["cntl"])
    fe.factor(0, 1e-8)
    yield fe, oracle_mod_big.csr()
    fe.close()


def test_batched_run_matches_oracle_per_source(tmp_path):
    """Four transmitters stepped as one block reproduce, per transmitter, the
    oracle's single-source run on the same (replayed) schedule to 1e-6."""
    from oracle import pyoracle
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    print(f"[gpu] batched run: {m.ntet} tets, {len(m.loops)} sources", flush=True)
    fe = gpu_from_model(m)
    rows = fe.run()
    keep = fe.checks()["keep"]
    fe.close()
    assert rows["fields"].shape[1:] == (4, 4, 9)
    # the same sources one at a time on the GPU, replaying the block schedule
    fe1 = gpu_from_model(m, loops=[2])
    fe1.set_replay(keep)
    r1 = fe1.run()
    fe1.close()
    np.testing.assert_array_equal(r1["t"], rows["t"])
    a, b = rows["fields"][:, 2, :, 5], r1["fields"][:, :, 5]
    sel = np.abs(b) > 1e-300
    rel = np.max(np.abs(a[sel] - b[sel]) / np.abs(b[sel]))
    print(f"[gpu] block vs single-RHS source 2: {rel:.3e}", flush=True)
    assert rel < 1e-9
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    for q in [0, 3]:
        root = str(tmp_path / f"src{q}")
        m.write(root, loop=q)
        o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
        o.assemble()
        o.set_replay(keep)
        ref = o.run()
        np.testing.assert_array_equal(ref["t"], rows["t"])
        a, b = rows["fields"][:, q, :, 5], ref["fields"][:, :, 5]
        sel = np.abs(b) > 1e-300
        rel = np.max(np.abs(a[sel] - b[sel]) / np.abs(b[sel]))
        print(f"[gpu] batched source {q} vs oracle: dBz/dt max rel diff {rel:.3e}", flush=True)
        assert rel < 1e-6


@pytest.mark.parametrize("loops", [[0, 1, 2, 3], [1]])
def test_graph_replay_matches_eager(loops):
    """Regular steps replayed from captured graphs (one per factor slot and
    arrangement of the history vectors) write the same rows, bit for bit, as
    the same run launched kernel by kernel; block and single-RHS paths."""
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    out = {}
    for graphs in (1, 0):
        print(f"[gpu] graph replay: sources {loops}, graphs={graphs}", flush=True)
        fe = gpu_from_model(m, loops=loops)
        fe.set_tuning("graphs", graphs)
        out[graphs] = (fe.run(), fe.checks())
        fe.close()
    (a, ca), (b, cb) = out[1], out[0]
    assert len(a["step"]) > m.sp_division + 3 * m.sp_double
    np.testing.assert_array_equal(ca["keep"], cb["keep"])
    for k in ("step", "t", "dt", "fields"):
        np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("loops", [[0, 1, 2, 3], [1]])
def test_overlapped_factor_matches_serial(loops):
    """The check's factorisation of P(2 dt), run on a second stream while the
    steps before the check go on, gives the rows, decisions and number of
    factorisations of the run that factors at the check itself."""
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    out = {}
    for ov in (1, 0):
        print(f"[gpu] overlapped factorisation: sources {loops}, overlap={ov}", flush=True)
        fe = gpu_from_model(m, loops=loops)
        fe.set_tuning("overlap", ov)
        out[ov] = (fe.run(), fe.checks(), fe.timers()["nfactor"])
        fe.close()
    (a, ca, na), (b, cb, nb) = out[1], out[0]
    assert len(ca["keep"]) >= 3 and na == nb
    np.testing.assert_array_equal(ca["keep"], cb["keep"])
    for k in ("step", "t", "dt", "fields"):
        np.testing.assert_array_equal(a[k], b[k])


def test_chained_rhs_matches_direct():
    """Regular steps of a block run form M (4 e2 - e3) as 4 M e2 - M e3 from
    the product kept from the step before (rhs_chain): the run through three
    doubling checks (after each of which the kept products go stale and are
    formed again) agrees with the direct product to rounding,
    with the same steps, times and decisions."""
    from temf3dt_amd import workloads
    m = workloads.multi_source_small_model()
    out = {}
    for ch in (1, 0):
        print(f"[gpu] chained right-hand side: rhs_chain={ch}", flush=True)
        fe = gpu_from_model(m)
        fe.set_tuning("rhs_chain", ch)
        out[ch] = (fe.run(), fe.checks())
        fe.close()
    (a, ca), (b, cb) = out[1], out[0]
    print(f"[gpu] check decisions {list(ca['keep'])}", flush=True)
    assert len(ca["keep"]) >= 3
    np.testing.assert_array_equal(ca["keep"], cb["keep"])
    for k in ("step", "t", "dt"):
        np.testing.assert_array_equal(a[k], b[k])
    fa, fb = a["fields"], b["fields"]
    scale = np.maximum(np.abs(fb).max(axis=(0, 1), keepdims=True), 1e-300)
    err = (np.abs(fa - fb) / scale).max(axis=(0, 1))  # per component (E, dB/dt)
    print(f"[gpu] chained vs direct: max difference per component {err} of its largest value", flush=True)
    # dB/dt tightly; E at receivers in the air carries its gradient part,
    # fixed only by the air conductivity, which amplifies rounding
    assert err[3:].max() < 1e-8, err
    assert err.max() < 1e-6, err


@pytest.mark.parametrize("knobs", [
    {"mt_units": 64, "mt_chmin": 16},                        # every long tile split into partials
    {"mt_units": 64, "mt_chmin": 16, "mt_ks_n": 4, "mt_ks_t": 2},
    {"mt_units": 100000, "mt_chmin": 4096, "mt_ks_n": 2, "mt_ks_t": 4},  # no splits
    {"mt_wide": 0},        # 32-row tiles on every level
    {"mt_wide": 100000, "mt_units": 64, "mt_chmin": 16},  # 64-row tiles on every level, split
])
def test_block_solve_tuning_knobs(gpu_mod_big_batch, knobs):
    """The tile-kernel knobs change only the schedule and the summation
    order: the block solve agrees with the default setting and keeps a
    backward error at FP64 level, including the partial-sum reductions that
    small meshes only reach with small chunks."""
    fe, oc = gpu_mod_big_batch
    rng = np.random.default_rng(7)
    B = rng.standard_normal((24, fe.nE))
    X0 = fe.solve(0, B)
    print(f"[gpu] tuning knobs {knobs}", flush=True)
    for k, v in knobs.items():
        fe.set_tuning(k, v)
    try:
        X = fe.solve(0, B)
        ms = fe.time_solve(0, 24, 2)
    finally:
        for k, v in {"mt_units": 4096, "mt_chmin": 256, "mt_ks_n": 2, "mt_ks_t": 2, "mt_wide": 64}.items():
            fe.set_tuning(k, v)
    assert ms > 0
    rel = np.linalg.norm(X - X0) / np.linalg.norm(X0)
    print(f"[gpu] knobs vs default: {rel:.3e}", flush=True)
    assert rel < 1e-10
    va = 2.0 * 1e-8 * oc["curl"] + 3.0 * oc["mass"]
    R = _upper_matvec(oc["ia"] - 1, oc["ja"] - 1, va, X.T) - B.T
    berr = np.abs(R).max() / (np.abs(va).max() * np.abs(X).max() + np.abs(B).max())
    assert berr < 1e-12


@pytest.mark.parametrize("band", [0, 1, 3])
def test_factor_update_order_is_bit_identical(gpu_mod_big_batch, band):
    """The trailing update's tile order (row by row, or bands of tile rows
    walked column by column) only changes which tiles run together: every
    tile sums the same terms in the same order, so the factor and the solve
    are bit-identical to those of the default order."""
    fe, _ = gpu_mod_big_batch
    rng = np.random.default_rng(11)
    B = rng.standard_normal((3, fe.nE))
    X0 = fe.solve(0, B)
    print(f"[gpu] factor update tile order: f_band={band}", flush=True)
    fe.set_tuning("f_band", band)
    try:
        fe.factor(0, 1e-8)
        X = fe.solve(0, B)
    finally:
        fe.set_tuning("f_band", 16)
        fe.factor(0, 1e-8)
    np.testing.assert_array_equal(X, X0)


def test_tuning_rejects_bad_values(gpu_mod_big_batch):
    from temf3dt_amd.fetd import FetdError
    fe, _ = gpu_mod_big_batch
    for k, v in [("mt_ks_n", 3), ("mt_chmin", 10), ("mt_units", 1), ("graphs", 2), ("overlap", -1), ("mt_wide", -1),
                 ("mt_fuse", 3), ("rhs_chain", 2), ("f_band", -1), ("f_band", 1025), ("keep_factors", -1), ("keep_states", 2), ("pinv", -1), ("s_tiles", 0), ("no_such_knob", 1)]:
        with pytest.raises(FetdError):
            fe.set_tuning(k, v)


def test_kernel_profile_counts(gpu_mod_big_batch):
    """HIP-event kernel timing: a block solve's tile launches are counted
    and timed, a single-RHS solve lands in its own category."""
    fe, _ = gpu_mod_big_batch
    fe.set_profiling(True)
    fe.time_solve(0, 8, 2)
    fe.time_solve(0, 1, 1)
    prof = fe.profile()
    fe.set_profiling(False)
    print(f"[gpu] profile {prof}", flush=True)
    assert prof["tile"][0] > 0 and prof["tile"][1] > 0
    assert prof["pull"][0] > 0
    assert prof["solve1"][0] > 0
    assert prof["factor_update"][0] == 0
