"""The Fortran driver's readers and writers (temf3dt_amd/fortran/fetd_io.f90),
run without the GPU through fetd_io_check: the MEDIT/.cntl readers agree
with the Python readers, and the .field writer reproduces the reference's
shipped .field from the oracle's rows byte for byte."""
import os
import subprocess

import numpy as np
import pytest

from temf3dt_amd import build as tfbuild
from temf3dt_amd import io as tfio

CHECK = tfbuild.FIO_CHECK

pytestmark = pytest.mark.skipif(not os.path.exists(CHECK), reason="Fortran programs not built (no amdflang)")


def _run(*args):
    r = subprocess.run([CHECK, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _write_rows(path, steps, times, fields, gates, o_f):
    nrec = fields.shape[1]
    with open(path, "w") as f:
        f.write(f"{nrec} {len(gates)}\n{' '.join(str(int(v)) for v in o_f)}\n")
        for g in gates:
            f.write(f"{float(g)!r}\n")
        f.write(f"{len(steps)}\n")
        for s, t, v in zip(steps, times, fields):
            f.write(f"{int(s)} {float(t)!r} " + " ".join(repr(float(x)) for x in v.reshape(-1)) + "\n")


def test_fortran_reader_matches_python(src_mod_big):
    out = _run("read", src_mod_big["root"])
    lines = [ln.split() for ln in out.strip().split("\n")]
    kv = {}
    for ln in lines:
        kv.setdefault(ln[0], []).append(ln[1:])
    mesh = tfio.read_mesh(src_mod_big["mesh"])
    c = tfio.read_cntl(src_mod_big["cntl"])
    nN, nT, nseg = (int(x) for x in kv["counts"][0])
    assert (nN, nT, nseg) == (len(mesh["xyz"]), len(mesh["tet_vex"]), len(mesh["seg_vex"]))
    assert float(kv["xyzsum"][0][0]) == pytest.approx(mesh["xyz"].sum(), rel=1e-12)  # summation order
    np.testing.assert_array_equal(np.array(kv["vfirst"][0], float), mesh["xyz"][0])
    np.testing.assert_array_equal(np.array(kv["vlast"][0], float), mesh["xyz"][-1])
    tv = mesh["tet_vex"].astype(np.int64)
    assert int(kv["tetsum"][0][0]) == int((tv * np.array([1, 2, 3, 4])).sum())
    assert int(kv["physum"][0][0]) == int(mesh["tet_phy"].astype(np.int64).sum())
    np.testing.assert_array_equal(np.array(kv["seg"], dtype=np.int64), mesh["seg_vex"])
    np.testing.assert_array_equal(np.array(kv["rec"], dtype=float), c["receivers"])
    np.testing.assert_array_equal(np.array(kv["dom"], dtype=float), np.array(c["domains"], float))
    st = kv["source"][0]
    assert (int(st[0]), int(st[1]), float(st[2]), float(st[3])) == (
        c["source_type"], c["pulse_type"], c["pulse_width"], c["current"])
    sc = kv["schedule"][0]
    assert [int(x) for x in sc[:4]] == [c["sp_mthd"], c["sp_division"], c["sp_double"], c["sp_dblesize"]]
    assert float(sc[4]) == c["time_maximum"]
    assert [int(x) for x in kv["output"][0]] == list(c["output"])


def test_fortran_writer_reproduces_shipped_field(src_mod_big, tmp_path):
    """The oracle's rows of src/mod_big written by the Fortran writer equal
    the reference's shipped .field byte for byte."""
    d = np.load(src_mod_big["oracle_rows"])
    inp = tmp_path / "rows.txt"
    _write_rows(inp, d["step"], d["t"], d["fields"], [], [0, 0, 0, 0, 0, 1, 0, 0, 0])
    _run("write", str(inp), str(tmp_path / "out"))
    with open(tmp_path / "out.field", "rb") as a, open(src_mod_big["ref_field"], "rb") as b:
        ours, ref = a.read().replace(b"\r\n", b"\n"), b.read().replace(b"\r\n", b"\n")
    assert ours == ref


def test_fortran_writer_gates_match_python(tmp_path):
    """Several receivers, all nine components and 40 gates: the Fortran
    .field/.field2 equal the Python writer's (same reference format and gate
    interpolation)."""
    rng = np.random.default_rng(3)
    nrow, nrec = 300, 3
    t = np.cumsum(np.concatenate([[0.0], np.full(nrow - 1, 1e-6) * (1 + np.arange(nrow - 1) // 50)]))
    fields = rng.standard_normal((nrow, nrec, 9)) * 10.0 ** rng.integers(-12, 3, (nrow, nrec, 9))
    fields[0] = 0.0
    gates = list(np.logspace(-6, np.log10(t[-1]) - 0.01, 40))
    o_f = [1] * 9
    inp = tmp_path / "rows.txt"
    _write_rows(inp, np.arange(nrow), t, fields, gates, o_f)
    _run("write", str(inp), str(tmp_path / "f"))
    tfio.write_field(str(tmp_path / "p.field"), np.arange(nrow), t, fields, o_f)
    tfio.write_field2(str(tmp_path / "p.field2"), t, fields, gates, o_f)
    for ext in ("field", "field2"):
        a = open(tmp_path / f"f.{ext}").read()
        b = open(tmp_path / f"p.{ext}").read()
        assert a == b, ext


def test_fortran_anisotropic_tensors_match_python(tmp_path):
    """The Fortran driver builds the same per-tet tensors (rotation with the
    reference's Rz entry, entrywise reciprocal permeability) as the Python
    host reader, bit for bit, from an 'anisotropic:' control file."""
    from temf3dt_amd import workloads
    m = workloads.aniso_small_model()
    root = str(tmp_path / "ani")
    m.write(root)
    out = _run("read", root)
    kv = {ln.split()[0]: ln.split()[1:] for ln in out.strip().split("\n")}
    mesh = tfio.read_mesh(root + ".mesh")
    c = tfio.read_cntl(root + ".cntl")
    s9, n9 = tfio.materials_tensor(mesh["tet_phy"], c["aniso"])
    for key, arr in (("tsig1", s9[0]), ("tnu1", n9[0]), ("tsigN", s9[-1]), ("tnuN", n9[-1])):
        np.testing.assert_array_equal(np.array(kv[key], float), arr.reshape(9))


def test_fortran_rescale_matches_python(src_small):
    """The small-loop variant's domain table (src_small/FETD_FWD.f90:314-328)."""
    out = _run("rescale", src_small["root"])
    kv = {}
    for ln in out.strip().split("\n"):
        kv.setdefault(ln.split()[0], []).append(ln.split()[1:])
    c = tfio.read_cntl(src_small["cntl"])
    doms, sp = tfio.rescaled_domains(c["domains"])
    assert float(kv["sp"][0][0]) == sp
    np.testing.assert_array_equal(np.array(kv["dom"], dtype=float), np.array(doms, float))


def test_fortran_writer_rescaled_matches_python(tmp_path):
    """Rows written with the small-loop variant's scaling v / sp**2 *
    pi**log10(sp): Fortran and Python writers agree byte for byte."""
    rng = np.random.default_rng(5)
    nrow, nrec = 120, 2
    t = np.cumsum(np.full(nrow, 3.3e-9))
    fields = rng.standard_normal((nrow, nrec, 9)) * 10.0 ** rng.integers(-14, 0, (nrow, nrec, 9))
    gates = list(np.logspace(-8, np.log10(t[-1]) - 0.01, 10))
    o_f = [0, 0, 0, 1, 1, 1, 0, 0, 1]
    inp = tmp_path / "rows.txt"
    _write_rows(inp, np.arange(nrow), t, fields, gates, o_f)
    _run("write", str(inp), str(tmp_path / "f"), repr(1000.0))
    tfio.write_field(str(tmp_path / "p.field"), np.arange(nrow), t, fields, o_f, 1000.0)
    tfio.write_field2(str(tmp_path / "p.field2"), t, fields, gates, o_f, 1000.0)
    for ext in ("field", "field2"):
        assert open(tmp_path / f"f.{ext}").read() == open(tmp_path / f"p.{ext}").read(), ext


def test_fortran_msh_slice_matches_python(src_small, tmp_path):
    """ROOT_i.msh from the Fortran driver's writer equals the Python writer's."""
    nT = len(tfio.msh_tet_tags(src_small["msh"]))
    rng = np.random.default_rng(11)
    f9 = rng.standard_normal((nT, 9)) * 10.0 ** rng.integers(-15, 2, (nT, 9))
    o_f = [1, 1, 0, 0, 0, 1, 0, 1, 1]
    inp = tmp_path / "vals.txt"
    with open(inp, "w") as f:
        f.write(f"{nT}\n{' '.join(map(str, o_f))}\n")
        f.write("".join(" ".join(repr(float(x)) for x in row) + "\n" for row in f9))
    _run("slice", src_small["msh"], str(inp), str(tmp_path / "f_1.msh"))
    tfio.write_msh_slice(src_small["msh"], str(tmp_path / "p_1.msh"), f9, o_f)
    with open(tmp_path / "f_1.msh", "rb") as a, open(tmp_path / "p_1.msh", "rb") as b:
        assert a.read() == b.read()
