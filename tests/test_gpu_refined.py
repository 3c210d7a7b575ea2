"""A locally refined mesh (newest-vertex bisection at the wire, graded
back to the grid's cells; temf3dt_amd/refine.py) through the whole path:
GPU against the oracle on the refined small model: rows and times
identical on the GPU's doubling schedule, the doubling checks' errors
within 1e-3 relative, dB/dt within 1e-6 relative."""
import os

import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu


def test_refined_small_model_matches_oracle(tmp_path):
    import dataclasses
    from oracle import pyoracle
    from temf3dt_amd import workloads
    m = dataclasses.replace(workloads.refined_small_model(), time_maximum=2e-4)
    print(f"[gpu] refined small model: {m.ntet} tets, {len(m.loops[0])} wire segments", flush=True)
    fe = gpu_from_model(m)
    rows = fe.run()
    ck = fe.checks()
    keep = ck["keep"]
    piv = fe.factor_pivots()
    fe.close()
    assert piv["negative"] == 0
    root = str(tmp_path / "m")
    m.write(root)
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    o.assemble()
    o.set_replay(keep)
    r = o.run()
    o.close()
    np.testing.assert_array_equal(rows["step"], r["step"])
    np.testing.assert_array_equal(rows["t"], r["t"])
    np.testing.assert_allclose(ck["err"], r["chk_err"], rtol=1e-3)  # the decisions' inputs
    a, b = rows["fields"][..., 3:6], r["fields"][..., 3:6]
    scale = np.abs(b).max()
    big = np.abs(b) > 1e-6 * scale
    rel = np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))
    print(f"[gpu] refined small model: {len(rows['t'])} rows, max rel dB/dt diff {rel:.3e}", flush=True)
    assert rel < 1e-6
    assert np.abs(a - b).max() < 1e-6 * scale
