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
