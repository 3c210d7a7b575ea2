import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def src_mod_big(tmp_path_factory):
    """The reference's shipped src/mod_big_sig_0.001 model, decompressed."""
    d = tmp_path_factory.mktemp("src_mod_big")
    g = os.path.join(GOLDEN, "src_mod_big")
    with gzip.open(os.path.join(g, "mod_big_sig_0.001.mesh.gz"), "rb") as f, open(d / "mod_big_sig_0.001.mesh",
                                                                                 "wb") as o:
        shutil.copyfileobj(f, o)
    shutil.copy(os.path.join(g, "mod_big_sig_0.001.cntl"), d / "mod_big_sig_0.001.cntl")
    shutil.copy(os.path.join(g, "mod_big_sig_0.001.field"), d / "ref.field")
    return dict(dir=str(d), root=str(d / "mod_big_sig_0.001"), mesh=str(d / "mod_big_sig_0.001.mesh"),
                cntl=str(d / "mod_big_sig_0.001.cntl"), ref_field=str(d / "ref.field"),
                oracle_rows=os.path.join(g, "oracle_rows.npz"))


@pytest.fixture(scope="session")
def src_small(tmp_path_factory):
    """The reference's shipped src_small/mod_small_sig_0.001 model (the
    small-loop variant with the conductivity rescaling), decompressed, with
    its gmsh .msh for the field-slice writer."""
    d = tmp_path_factory.mktemp("src_small")
    g = os.path.join(GOLDEN, "src_small")
    for ext in ("mesh", "msh"):
        with gzip.open(os.path.join(g, f"mod_small_sig_0.001.{ext}.gz"), "rb") as f, \
                open(d / f"mod_small_sig_0.001.{ext}", "wb") as o:
            shutil.copyfileobj(f, o)
    shutil.copy(os.path.join(g, "mod_small_sig_0.001.cntl"), d / "mod_small_sig_0.001.cntl")
    shutil.copy(os.path.join(g, "mod_small_sig_0.001.field"), d / "ref.field")
    return dict(dir=str(d), root=str(d / "mod_small_sig_0.001"), mesh=str(d / "mod_small_sig_0.001.mesh"),
                msh=str(d / "mod_small_sig_0.001.msh"), cntl=str(d / "mod_small_sig_0.001.cntl"),
                ref_field=str(d / "ref.field"))


def replay_from_field(path, sp_double):
    """Keep (1) / change (0) outcome of every doubling check, read off a
    .field file's time column: at a check step the time advances by the old
    dt on a keep and by twice it on a change (src/FETD_FWD.f90:1380-1417)."""
    rows = np.loadtxt(path)
    st, t = rows[:, 0].astype(int), rows[:, 1]
    keep = []
    for i in range(2, len(st)):
        if st[i] % sp_double == 0:
            keep.append(1 if (t[i] - t[i - 1]) / (t[i - 1] - t[i - 2]) < 1.5 else 0)
    return keep


@pytest.fixture(scope="session")
def oracle_mod_big(src_mod_big):
    from oracle import pyoracle
    o = pyoracle.Oracle(src_mod_big["mesh"], src_mod_big["cntl"])
    o.assemble()
    return o


def gpu_from_model(m, loops=None, device=0):
    """Fetd context for a workloads.Model; `loops` selects transmitters
    (default: all of them, run as one block)."""
    from temf3dt_amd import fetd
    loops = list(range(len(m.loops))) if loops is None else list(loops)
    fe = fetd.Fetd(device)
    fe.set_mesh(m.xyz, m.tet_vex)
    sig, imu = m.sigma_inv_mu()
    fe.set_materials(sig, imu)
    if len(loops) == 1:
        fe.set_source(m.loops[loops[0]], *m.source)
    else:
        fe.set_sources([m.loops[q] for q in loops], *m.source)
    fe.set_receivers(m.receivers, m.rec_charac)
    fe.set_schedule(m.sp_division, m.sp_double, m.sp_dblesize, m.time_maximum)
    fe.assemble()
    return fe
