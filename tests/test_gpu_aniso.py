"""Anisotropic media (tensor conductivity and permeability per domain) on
the GPU against the oracle: assembled matrices and the full run."""
import os

import numpy as np
import pytest

from temf3dt_amd import fetd

pytestmark = pytest.mark.gpu


def test_anisotropic_run_matches_oracle(tmp_path):
    from oracle import pyoracle
    from temf3dt_amd import workloads
    m = workloads.aniso_small_model()
    root = str(tmp_path / "ani")
    m.write(root)
    print(f"[gpu] anisotropic model: {m.ntet} tets", flush=True)
    fe, _, _ = fetd.setup_from_files(root + ".mesh", root + ".cntl")
    M, K = fe.matrices()
    rows = fe.run()
    keep = fe.checks()["keep"]
    fe.close()
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    o.assemble()
    oc = o.csr()
    dm = np.abs(M - oc["mass"]).max() / np.abs(oc["mass"]).max()
    dk = np.abs(K - oc["curl"]).max() / np.abs(oc["curl"]).max()
    print(f"[gpu] anisotropic assembly: mass {dm:.2e} curl {dk:.2e}", flush=True)
    assert dm < 1e-14 and dk < 1e-14
    o.set_replay(keep)
    r = o.run()
    o.close()
    np.testing.assert_array_equal(rows["step"], r["step"])
    a, b = rows["fields"][:, :, 5], r["fields"][:, :, 5]
    scale = np.abs(b).max()
    big = np.abs(b) > 1e-6 * scale
    rel = np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))
    print(f"[gpu] anisotropic run: {len(rows['t'])} rows, max rel dBz/dt diff {rel:.3e}", flush=True)
    assert rel < 1e-6
