"""Per-tet field slices (vtkout) and the small-loop variant on the GPU,
against the oracle and the reference's shipped src_small/mod_small_sig_0.001
files."""
import os
import shutil

import numpy as np
import pytest

from conftest import replay_from_field

pytestmark = pytest.mark.gpu


def _oracle(src_small, keep, times, max_steps):
    from oracle import pyoracle
    pyoracle.set_threads(min(16, os.cpu_count() or 8))
    o = pyoracle.Oracle(src_small["mesh"], src_small["cntl"], sig_rescale=True)
    o.assemble()
    o.set_replay(keep)
    o.set_slices(times)
    r = o.run(max_steps=max_steps)
    return o, r


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_slices_match_oracle(src_small):
    from temf3dt_amd import fetd
    keep = replay_from_field(src_small["ref_field"], 32)
    ref = np.loadtxt(src_small["ref_field"])
    times = [0.0, ref[50, 1], 0.5 * (ref[120, 1] + ref[121, 1]), ref[199, 1]]
    print(f"[gpu] src_small slices at {times}, 200 steps", flush=True)
    fe, m, c = fetd.setup_from_files(src_small["mesh"], src_small["cntl"], sig_rescale=True)
    fe.set_replay(keep)
    fe.set_slices(times)
    rows = fe.run(max_steps=200)
    got = [fe.slice(i) for i in range(len(times))]
    fe.close()
    o, orow = _oracle(src_small, keep, times, 200)
    np.testing.assert_array_equal(rows["step"], orow["step"])
    np.testing.assert_array_equal(rows["t"], orow["t"])
    earth = m["tet_phy"] != 11
    for i, (g, gs, gt) in enumerate(got):
        f, s, t = o.slice(i)
        assert (gs, gt) == (s, t), (i, gs, gt, s, t)
        if i == 0:
            assert np.all(g == 0.0) and np.all(f == 0.0)
            continue
        rd = [_rel(g[:, k], f[:, k]) for k in range(3, 9)]
        re = [_rel(g[earth, k], f[earth, k]) for k in range(3)]
        print(f"[gpu] slice {i} (step {s}): curl E, B max diff / max {np.array2string(np.array(rd), precision=2)}; "
              f"E in the earth {np.array2string(np.array(re), precision=2)}", flush=True)
        assert max(rd) < 1e-6
        assert max(re) < 1e-6


def test_small_variant_end_to_end(src_small, tmp_path):
    """run_root with the conductivity rescaling and the shipped schedule:
    the .field matches the reference's 5-digit file and every vtkout slice
    is written as ROOT_i.msh with the oracle's |dBz/dt| per tet."""
    from temf3dt_amd import fetd, io as tfio
    root = str(tmp_path / "m")
    shutil.copy(src_small["mesh"], root + ".mesh")
    shutil.copy(src_small["msh"], root + ".msh")
    times = [1e-6, 1e-4, 1e-2, 1.0]  # the last is never reached
    cntl = open(src_small["cntl"]).read().replace("vtkout: 0", "vtkout: 4\n" + "\n".join(map(str, times)))
    open(root + ".cntl", "w").write(cntl)
    keep = replay_from_field(src_small["ref_field"], 32)
    print("[gpu] run_root src_small, rescaled, shipped schedule", flush=True)
    rows = fetd.run_root(root, sig_rescale=True, replay=keep)
    assert len(rows["step"]) == 923
    ref = np.loadtxt(src_small["ref_field"])
    out = np.loadtxt(root + ".field")
    np.testing.assert_array_equal(out[:, :2], ref[:, :2])
    sel = ref[:, 2] != 0
    rel = np.abs(out[sel, 2] - ref[sel, 2]) / np.abs(ref[sel, 2])
    same = sum(a == b for a, b in zip(open(root + ".field"), open(src_small["ref_field"])))
    print(f"[gpu] .field: {same}/923 lines identical, max rel diff {rel.max():.2e}", flush=True)
    assert rel.max() < 6e-5  # 5 significant digits
    o, _ = _oracle(src_small, keep, times, 0)
    tags = tfio.msh_tet_tags(root + ".msh")
    for i in range(3):
        path = tfio.slice_path(root, i + 1)
        lines = open(path).read().splitlines()[-len(tags) - 1:-1]
        vals = np.array([float(ln[10:]) for ln in lines])
        f, _, _ = o.slice(i)
        r = np.abs(vals - np.abs(f[:, 5])).max() / np.abs(f[:, 5]).max()
        print(f"[gpu] {os.path.basename(path)}: |dBz/dt| max diff / max {r:.2e}", flush=True)
        assert r < 1e-6
    assert not os.path.exists(tfio.slice_path(root, 4))
