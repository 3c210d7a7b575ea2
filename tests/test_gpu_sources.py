"""Source waveforms beyond step-on: impulse (pulse type 2) and the two-pass
step-off (pulse type 3: a negated step-on run to the end time, whose final
field starts the second run; src/FETD_FWD.f90:1212-1217,1295-1301,1453-1461),
GPU against the oracle on the same model."""
import dataclasses
import os

import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pulse_type", [2, 3])
def test_pulse_type_matches_oracle(pulse_type, tmp_path):
    from oracle import pyoracle
    from temf3dt_amd import workloads
    m = workloads.small_model()
    st, _, t0, cur = m.source
    m = dataclasses.replace(m, source=(st, pulse_type, t0, cur), time_maximum=2e-4)
    print(f"[gpu] pulse type {pulse_type}: {m.ntet} tets", flush=True)
    fe = gpu_from_model(m)
    rows = fe.run()
    keep = fe.checks()["keep"]
    fe.close()
    root = str(tmp_path / "m")
    m.write(root)
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    o.assemble()
    o.set_replay(keep)
    r = o.run()
    o.close()
    np.testing.assert_array_equal(rows["step"], r["step"])
    np.testing.assert_allclose(rows["t"], r["t"], rtol=0, atol=0)
    a, b = rows["fields"][:, :, 5], r["fields"][:, :, 5]
    scale = np.abs(b).max()
    big = np.abs(b) > 1e-6 * scale
    rel = np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))
    print(f"[gpu] pulse type {pulse_type}: {len(rows['t'])} rows, max rel dBz/dt diff {rel:.3e} "
          f"(|b| > 1e-6 max)", flush=True)
    assert rel < 1e-6
    assert np.abs(a - b).max() < 1e-6 * scale


def test_stepoff_first_pass_ending_on_changed_check(tmp_path):
    """Step-off whose first pass ends on a doubling check that changed dt:
    the second pass must start from the reference's e1 (the regular step
    before the check, src/FETD_FWD.f90:1459), not the check's e_veri_1.
    time_maximum is set to the new time of the first changed check, so the
    first pass's loop ends right there."""
    from oracle import pyoracle
    from temf3dt_amd import workloads
    m = workloads.small_model()
    st, _, t0, cur = m.source
    probe = dataclasses.replace(m, time_maximum=2e-4)
    fe = gpu_from_model(probe)
    rows = fe.run()
    ck = fe.checks()
    fe.close()
    chg = [s for s, k in zip(ck["step"], ck["keep"]) if k == 0 and s > 0]
    assert chg
    r = int(np.nonzero(rows["step"] == chg[0])[0][0])
    m = dataclasses.replace(m, source=(st, 3, t0, cur), time_maximum=float(rows["t"][r]))
    print(f"[gpu] step-off ending on the changed check at step {chg[0]}, t_max {m.time_maximum:.6e}", flush=True)
    fe = gpu_from_model(m)
    got = fe.run()
    keep = fe.checks()["keep"]
    fe.close()
    root = str(tmp_path / "m")
    m.write(root)
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
    o.assemble()
    o.set_replay(keep)
    ref = o.run()
    o.close()
    np.testing.assert_array_equal(got["step"], ref["step"])
    np.testing.assert_array_equal(got["t"], ref["t"])
    # row 0 carries E of the handed-over field
    e0, r0 = got["fields"][0, :, 0:3], ref["fields"][0, :, 0:3]
    assert np.abs(e0 - r0).max() <= 1e-9 * np.abs(r0).max()
    a, b = got["fields"][:, :, 5], ref["fields"][:, :, 5]
    scale = np.abs(b).max()
    big = np.abs(b) > 1e-6 * scale
    rel = np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))
    print(f"[gpu] step-off (changed-check hand-off): {len(got['t'])} rows, max rel dBz/dt diff {rel:.3e}", flush=True)
    assert rel < 1e-6
