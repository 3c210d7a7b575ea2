import gzip
import os
import shutil
import sys

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
def oracle_mod_big(src_mod_big):
    from oracle import pyoracle
    o = pyoracle.Oracle(src_mod_big["mesh"], src_mod_big["cntl"])
    o.assemble()
    return o
