"""Write full-precision oracle vectors for the shipped src/mod_big_sig_0.001
model (tests/golden/src_mod_big): every .field row (step, t, dt, 9 fields)
and the doubling-check log.  The oracle's own .field output for this model
is byte-identical to the reference's shipped mod_big_sig_0.001.field, which
pins these vectors; they let GPU tests compare at 1e-6 without re-running
the CPU restatement.

Usage: python tests/golden/make_oracle_rows.py
"""
import gzip
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import pyoracle  # noqa: E402


def main():
    d = os.path.join(HERE, "src_mod_big")
    with tempfile.TemporaryDirectory() as tmp:
        mesh = os.path.join(tmp, "m.mesh")
        with gzip.open(os.path.join(d, "mod_big_sig_0.001.mesh.gz"), "rb") as f, open(mesh, "wb") as g:
            shutil.copyfileobj(f, g)
        pyoracle.set_threads(os.cpu_count() or 8)
        o = pyoracle.Oracle(mesh, os.path.join(d, "mod_big_sig_0.001.cntl"))
        o.assemble()
        field = os.path.join(tmp, "o.field")
        r = o.run(field=field)
        ref = open(os.path.join(d, "mod_big_sig_0.001.field")).read()
        got = open(field).read()
        if ref != got:
            raise SystemExit("oracle .field differs from the reference's shipped file")
    np.savez_compressed(os.path.join(d, "oracle_rows.npz"), step=r["step"], t=r["t"], dt=r["dt"],
                        fields=r["fields"], chk_step=r["chk_step"], chk_err=r["chk_err"], chk_keep=r["chk_keep"])
    print("rows", len(r["step"]), "checks", len(r["chk_step"]), "factor nnz", r["factor_nnz"])


if __name__ == "__main__":
    main()
