"""Reference file formats: MEDIT .mesh, .cntl control files, .field/.field2.

Host-side mirror of the reference's input/output code:
  * read_mesh   -- model_mesh's reader, src/FETD_FWD.f90:126-167
  * read_cntl   -- model_ctl's parser (parseCode grammar,
                   src/FETD_UTILITY.f90:678-731; keys src/FETD_FWD.f90:313-497)
  * FieldWriter -- solution_output's .field/.field2 rows,
                   src/FETD_FWD.f90:1602-1641 (format "(i6,e15.5,2x,R(m e15.5,2x))")
  * write_msh_slice -- msh_output's per-tet field slices appended to a copy
                   of ROOT.msh, src/FETD_FWD.f90:1713-1864
  * sig_parameter -- the small-loop variant's conductivity rescaling,
                   src_small/FETD_FWD.f90:314-328,1526,1540
"""
import gzip
import math
import shutil

import numpy as np

MU0 = math.pi * 4.0e-7            # src/FETD_FWD.f90:15
EPS0 = 8.8541878176e-12           # src/FETD_FWD.f90:16


def _open(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


def _fnum(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def read_mesh(path):
    """MEDIT reader: sections keyed by the first five characters of the
    first token (Verti/Edges/Trian/Tetra).  Edges with a non-zero physical
    tag are the source wire segments, in file order."""
    with _open(path) as f:
        lines = f.read().split("\n")
    out = {}
    i = 0
    n = len(lines)

    def block(count, ncol):
        nonlocal i
        rows = []
        while len(rows) < count:
            i += 1
            t = lines[i].split()
            if t:
                rows.append(t[:ncol])
        return rows

    while i < n:
        t = lines[i].split()
        if t:
            w = t[0][:5]
            if w in ("Verti", "Edges", "Trian", "Tetra"):
                i += 1
                while not lines[i].split():
                    i += 1
                cnt = int(lines[i].split()[0])
                if w == "Verti":
                    rows = block(cnt, 3)
                    out["xyz"] = np.array([[_fnum(v) for v in r] for r in rows], dtype=np.float64).reshape(-1, 3)
                elif w == "Edges":
                    rows = np.array(block(cnt, 3), dtype=np.int64).reshape(-1, 3)
                    keep = rows[:, 2] != 0
                    out["seg_vex"] = rows[keep, :2].astype(np.int32)
                    out["seg_phy"] = rows[keep, 2].astype(np.int32)
                elif w == "Trian":
                    block(cnt, 3)
                else:
                    rows = np.array(block(cnt, 5), dtype=np.int64).reshape(-1, 5)
                    out["tet_vex"] = np.ascontiguousarray(rows[:, :4].astype(np.int32))
                    out["tet_phy"] = rows[:, 4].astype(np.int32)
        i += 1
    if "xyz" not in out or "tet_vex" not in out:
        raise ValueError(f"{path}: mesh has no Vertices or Tetrahedra")
    out.setdefault("seg_vex", np.zeros((0, 2), np.int32))
    out.setdefault("seg_phy", np.zeros((0,), np.int32))
    return out


def _parse_code(line):
    """parseCode: returns (lower-cased key, value) or None."""
    s = line.replace("\t", " ")
    st = s.lstrip(" ")
    if not st or st[0] in "%!":
        return None
    if ":" not in s:
        return None
    ci = s.index(":")
    lead = len(s) - len(st)
    if ci < lead:
        return None
    key = s[lead:ci].rstrip(" ").lower()
    val = s[ci + 1:].lstrip(" ")
    for c in "%!":
        k = val.find(c)
        if k >= 0:
            val = val[:k]
    return key, val


class _Records:
    def __init__(self, lines):
        self.lines = lines
        self.i = 0

    def next(self):
        if self.i >= len(self.lines):
            return None
        ln = self.lines[self.i]
        self.i += 1
        return ln

    def nums(self, k):
        """Fortran list-directed read of k values from the next record(s)."""
        got = []
        while len(got) < k:
            ln = self.next()
            if ln is None:
                raise ValueError("truncated control file")
            for tok in ln.replace(",", " ").split():
                if len(got) < k:
                    try:
                        got.append(_fnum(tok))
                    except ValueError:
                        break
        return got


def read_cntl(path):
    """model_ctl's keys: receivers, isotropic, boundary condition, source,
    sp_mthd, sp_division, time_maximum, sp_double, sp_dblesize, output,
    vtkout, time_points, rec_charac.  Unknown keys are ignored."""
    with open(path, "r") as f:
        rec = _Records(f.read().replace("\r", "").split("\n"))
    c = dict(rec_charac=10.0, receivers=None, domains=None, bound_cond=0, src_num=0, sp_mthd=0,
             sp_division=0, time_maximum=0.0, sp_double=0, sp_dblesize=0, output=[0] * 9,
             time_points=[], vtk_times=[])
    while True:
        ln = rec.next()
        if ln is None:
            break
        pc = _parse_code(ln)
        if pc is None:
            continue
        key, val = pc
        first = val.split()[0] if val.split() else "0"
        if key == "time_points":
            c["time_points"] = [rec.nums(1)[0] for _ in range(int(_fnum(first)))]
        elif key == "rec_charac":
            c["rec_charac"] = _fnum(first)
        elif key == "receivers":
            c["receivers"] = np.array([rec.nums(3) for _ in range(int(_fnum(first)))], dtype=np.float64)
        elif key == "isotropic":
            c["domains"] = [rec.nums(4) for _ in range(int(_fnum(first)))]
        elif key == "anisotropic":
            # per domain three records (conductivity, permittivity,
            # permeability): "domain v1 v2 v3 angle1 angle2 angle3"
            doms = []
            for _ in range(int(_fnum(first))):
                recs = [rec.nums(7) for _ in range(3)]
                doms.append((int(recs[0][0]), recs[0][1:4], recs[0][4:7], recs[1][1:4], recs[1][4:7],
                             recs[2][1:4], recs[2][4:7]))
            c["aniso"] = doms
            c["domains"] = [(d[0], d[1][0], d[3][0], d[5][0]) for d in doms]
        elif key == "boundary condition":
            c["bound_cond"] = int(_fnum(first))
        elif key == "source":
            c["src_num"] = int(_fnum(first))
            st, pt, t0, cur = rec.nums(4)
            c["source_type"], c["pulse_type"], c["pulse_width"], c["current"] = int(st), int(pt), t0, cur
        elif key == "sp_mthd":
            c["sp_mthd"] = int(_fnum(first))
        elif key == "sp_division":
            c["sp_division"] = int(_fnum(first))
        elif key == "time_maximum":
            c["time_maximum"] = _fnum(first)
        elif key == "sp_double":
            c["sp_double"] = int(_fnum(first))
        elif key == "sp_dblesize":
            c["sp_dblesize"] = int(_fnum(first))
        elif key == "output":
            c["output"] = [int(v) for v in rec.nums(9)]
        elif key == "vtkout":
            nv = int(_fnum(first))
            c["vtk_times"] = [rec.nums(1)[0] for _ in range(nv)]
    # required keys (src/FETD_FWD.f90:490-497)
    checks = [(c["receivers"] is None, "cmod%rec_num error"), (c["domains"] is None, "cmod%dom_num error"),
              (c["src_num"] == 0, "cmod%src_num error"), (c["sp_mthd"] == 0, "cmod%step_mthd error"),
              (c["sp_division"] == 0, "cmod%step_dvsi error"),
              (abs(c["time_maximum"]) < 1e-20, "cmod%time_maxm error"),
              (c["sp_double"] == 0, "cmod%step_dble error"),
              (not any(v == 1 for v in c["output"]), "cmod%o_f error")]
    for bad, msg in checks:
        if bad:
            raise ValueError(msg)
    return c


def sig_parameter(domains):
    """The small-loop variant's scale sp (src_small/FETD_FWD.f90:314-323):
    1 / sigma of the last domain when there are several (the branch written
    to take the smallest sigma is unreachable), else 1."""
    sp = 1.0
    for i, d in enumerate(domains):
        if i > 0:
            sp = float(d[1])
    return 1.0 / sp


def rescaled_domains(domains):
    """Domain table of the small-loop variant: sigma of domains 2..n times sp
    (src_small/FETD_FWD.f90:324-328); returns (domains, sp)."""
    sp = sig_parameter(domains)
    out = [tuple(d) if i == 0 else (d[0], float(d[1]) * sp) + tuple(d[2:]) for i, d in enumerate(domains)]
    return out, sp


def output_scale(values, sp):
    """Written values of the small-loop variant: v / sp**2 * pi**log10(sp)
    (src_small/FETD_FWD.f90:1526,1540); sp = None leaves them as they are."""
    if sp is None:
        return values
    return np.asarray(values) / sp ** 2.0 * math.pi ** math.log10(sp)


def materials(tet_phy, domains, sig_rescale=False):
    """Per-tet sigma and 1/mu from the isotropic domain table
    (src/FETD_FWD.f90:347-364); sig_rescale applies the small-loop variant's
    scaling of domains 2..n."""
    if sig_rescale:
        domains, _ = rescaled_domains(domains)
    sigma = np.empty(len(tet_phy), np.float64)
    inv_mu = np.empty(len(tet_phy), np.float64)
    found = np.zeros(len(tet_phy), bool)
    for dom, sig, _per, mag in domains:
        m = tet_phy == int(dom)
        sigma[m] = sig
        inv_mu[m] = 1.0 / (mag * MU0)
        found |= m
    if not found.all():
        raise ValueError("error - physical domain not found!")
    return sigma, inv_mu


def _mat3(a, b):
    """3x3 product with sums in index order (matches the restated Fortran)."""
    return [[a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] for j in range(3)] for i in range(3)]


def aniso_tensor(values, angles_deg, scale=1.0):
    """model_ctl's tensor (src/FETD_FWD.f90:381-408): R diag(v) R^T with
    R = Rx Ry Rz from three angles in degrees; as in the reference, the (2,2)
    entry of Rz takes the cosine of the second angle.  Row-major 3x3 list."""
    d2r = math.pi / 180.0
    a1, a2, a3 = (float(x) * d2r for x in angles_deg)
    dx = [[1.0, 0.0, 0.0], [0.0, math.cos(a1), -math.sin(a1)], [0.0, math.sin(a1), math.cos(a1)]]
    dy = [[math.cos(a2), 0.0, math.sin(a2)], [0.0, 1.0, 0.0], [-math.sin(a2), 0.0, math.cos(a2)]]
    dz = [[math.cos(a3), -math.sin(a3), 0.0], [math.sin(a3), math.cos(a2), 0.0], [0.0, 0.0, 1.0]]
    dm = _mat3(_mat3(dx, dy), dz)
    pm = [[float(values[0]), 0.0, 0.0], [0.0, float(values[1]), 0.0], [0.0, 0.0, float(values[2])]]
    dmt = [[dm[j][i] for j in range(3)] for i in range(3)]
    a = _mat3(_mat3(dm, pm), dmt)
    return [[a[i][j] * scale for j in range(3)] for i in range(3)]


def materials_tensor(tet_phy, aniso):
    """Per-tet conductivity tensors and curl weights [nT][3][3] from the
    anisotropic domain table: the curl weight is the entrywise reciprocal of
    the permeability tensor, zero entries kept (src/FETD_FWD.f90:621-633)."""
    sig9 = np.empty((len(tet_phy), 3, 3))
    nu9 = np.empty((len(tet_phy), 3, 3))
    found = np.zeros(len(tet_phy), bool)
    for dom, sv, sa, _pv, _pa, mv, ma in aniso:
        S = np.array(aniso_tensor(sv, sa))
        Mg = np.array(aniso_tensor(mv, ma, MU0))
        Nu = np.where(Mg == 0.0, 0.0, 1.0 / np.where(Mg == 0.0, 1.0, Mg))
        sel = tet_phy == int(dom)
        sig9[sel] = S
        nu9[sel] = Nu
        found |= sel
    if not found.all():
        raise ValueError("error - physical domain not found!")
    return sig9, nu9


def fort_e(v, w=15, d=5):
    """Fortran Ew.d edit descriptor: [-]0.dddddE+xx right-justified."""
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(v):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        sign = "-" if v < 0 else ""
        if abs(ex) <= 99:
            s = f"{sign}0.{digits}E{'-' if ex < 0 else '+'}{abs(ex):02d}"
        else:
            s = f"{sign}0.{digits}{'-' if ex < 0 else '+'}{abs(ex):03d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def format_row(step, t, vals, o_f):
    """One .field row: i6, e15.5, 2x, then per receiver the selected
    components (e15.5 each) followed by 2x between receivers."""
    parts = [f"{step:6d}", fort_e(t), "  "]
    nrec = len(vals)
    for r in range(nrec):
        for k in range(9):
            if o_f[k]:
                parts.append(fort_e(vals[r][k]))
        if r < nrec - 1:
            parts.append("  ")
    return "".join(parts)


def write_field(path, steps, times, fields, o_f, sp=None):
    """fields: [rows][R][9]; sp: the small-loop variant's scale (or None)"""
    with open(path, "w") as f:
        for s, t, v in zip(steps, times, fields):
            f.write(format_row(int(s), float(t), output_scale(v, sp), o_f) + "\n")


def write_field2(path, times, fields, gates, o_f, sp=None):
    """.field2 rows at the time_points gates: the first output row with
    t >= gate is interpolated linearly with the previous row
    (src/FETD_FWD.f90:1625-1641, interp_lin src/FETD_UTILITY.f90:21-31)."""
    ok = [False] * len(gates)
    lines = []
    for r in range(len(times)):
        t = times[r]
        for g, tg in enumerate(gates):
            if t >= tg and not ok[g]:
                if r == 0:
                    break
                x1, x2 = times[r - 1], t
                y1, y2 = fields[r - 1], fields[r]
                a = (y1 - y2) / (x1 - x2)
                b = y1 - a * x1
                lines.append(format_row(g + 1, tg, output_scale(a * tg + b, sp), o_f))
                ok[g] = True
    with open(path, "w") as f:
        for ln in lines:
            f.write(ln + "\n")


# ---- per-tet field slices (msh_output, src/FETD_FWD.f90:1713-1864) -----------
SLICE_NAMES = ["Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Bx", "By", "Bz"]


def msh_tet_tags(path):
    """Element tags of the tetrahedra (element type 4) of a gmsh 4.x .msh, in
    file order: the $Elements walk of msh_output (src/FETD_FWD.f90:1744-1788),
    which reads each block's "dim tag type count" header and the first number
    of every element line."""
    with open(path) as f:
        for line in f:
            if line[:9] == "$Elements":
                break
        else:
            raise ValueError("cell number not found")
        nblock = int(f.readline().split()[0])
        tags = []
        for _ in range(nblock):
            _dim, _ent, etype, cnt = (int(x) for x in f.readline().split()[:4])
            for _ in range(cnt):
                tag = int(f.readline().split()[0])
                if etype == 4:
                    tags.append(tag)
    return np.array(tags, np.int64)


def slice_path(root, i):
    """ROOT_i.msh for slice i (1-based), as msh_output names it."""
    return f"{root}_{i}.msh"


def write_msh_slice(src_msh, dst, fields9, o_f):
    """A copy of src_msh with one $ElementData section per selected output
    component: the component's absolute value per tetrahedron, "(i10,e16.8)"
    rows after the gmsh header lines (src/FETD_FWD.f90:1789-1862).  fields9
    is [nT][9] (E, curl E, B) in the mesh's tet order."""
    tags = msh_tet_tags(src_msh)
    f9 = np.asarray(fields9)
    if len(tags) != len(f9):
        raise ValueError(f"{src_msh} has {len(tags)} tetrahedra, the fields {len(f9)}")
    shutil.copyfile(src_msh, dst)
    with open(dst, "a") as f:
        for k in range(9):
            if not o_f[k]:
                continue
            f.write("$ElementData\n1\n")
            f.write(f'"{SLICE_NAMES[k]} data"\n')
            f.write("1\n0.0\n3\n0\n1\n")
            f.write(f"{len(tags):12d}\n")
            col = np.abs(f9[:, k])
            f.write("".join(f"{int(t):10d}{fort_e(float(v), 16, 8)}\n" for t, v in zip(tags, col)))
            f.write("$EndElementData\n")
