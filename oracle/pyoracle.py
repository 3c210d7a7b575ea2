"""ctypes wrapper of liboracle.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (temf3dt_amd/).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.fo_load.restype = _P
        L.fo_load.argtypes = [ctypes.c_char_p, _I]
        L.fo_load_files.restype = _P
        L.fo_load_files.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I]
        L.fo_load_files_variant.restype = _P
        L.fo_load_files_variant.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I, _I]
        L.fo_sig_parameter.restype = ctypes.c_double
        L.fo_sig_parameter.argtypes = [_P]
        L.fo_set_slices.argtypes = [_P, _I, _P]
        L.fo_get_slice.argtypes = [_P, _I, _P, _P, _P]
        L.fo_tet_fields.argtypes = [_P, _P, _P]
        L.fo_error.restype = ctypes.c_char_p
        L.fo_factor_nnz.restype = ctypes.c_longlong
        L.fo_factor_nnz.argtypes = [_P]
        L.fo_det_lu.restype = _D
        L.fo_det_lu.argtypes = [_P, _I]
        L.fo_source_didt.restype = _D
        L.fo_source_didt.argtypes = [_I, _I, _D, _D, _D]
        for n in ["fo_free", "fo_assemble", "fo_num_rows", "fo_num_checks", "fo_num_factorizations",
                  "fo_pardiso_available"]:
            getattr(L, n).argtypes = [_P] if n != "fo_pardiso_available" else []
        L.fo_run.argtypes = [_P, ctypes.c_char_p, ctypes.c_char_p, _I]
        L.fo_set_threads.argtypes = [_I]
        L.fo_pardiso_solve_once.argtypes = [_I, _P, _P, _P, _P, _P, _I]
        L.fo_get_tensors.argtypes = [_P, _P, _P]
        L.fo_gate_rows.argtypes = [_I, _I, _P, _P, _I, _P, _P, _P]
        L.fo_gate_rows.restype = None
        L.fo_cpu_costs.argtypes = [_P, _D, _I, _P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


class Oracle:
    """One model in the CPU restatement."""

    def __init__(self, mesh_path, cntl_path, verbose=0, sig_rescale=False):
        L = lib()
        m = L.fo_load_files_variant(mesh_path.encode(), cntl_path.encode() if cntl_path else None, verbose,
                                    1 if sig_rescale else 0)
        if not m:
            raise RuntimeError(L.fo_error().decode())
        self.L = L
        self.m = _P(m)
        c = [_I() for _ in range(6)]
        L.fo_counts(self.m, *[ctypes.byref(x) for x in c])
        self.nN, self.nT, self.nE, self.nnz, self.nR, self.nseg = [x.value for x in c]

    def close(self):
        if self.m:
            self.L.fo_free(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self):
        if self.L.fo_assemble(self.m) != 0:
            raise RuntimeError(self.L.fo_error().decode())
        c = [_I() for _ in range(6)]
        self.L.fo_counts(self.m, *[ctypes.byref(x) for x in c])
        self.nnz = c[3].value

    def topology(self):
        edg = np.empty((self.nE, 2), np.int32)
        te = np.empty((self.nT, 6), np.int32)
        td = np.empty((self.nT, 6), np.int32)
        el = np.empty(self.nE)
        self.L.fo_get_topology(self.m, _p(edg), _p(te), _p(td), _p(el))
        return dict(edg_vex=edg, tet_edg=te, tet_drc=td, edg_len=el)

    def geometry(self):
        a, b, c, d = [np.empty((self.nT, 4)) for _ in range(4)]
        vol = np.empty(self.nT)
        tl = np.empty((self.nT, 6))
        self.L.fo_get_geometry(self.m, _p(a), _p(b), _p(c), _p(d), _p(vol), _p(tl))
        return dict(a=a, b=b, c=c, d=d, vol=vol, tet_len=tl)

    def csr(self):
        ia = np.empty(self.nE + 1, np.int32)
        ja = np.empty(self.nnz, np.int32)
        ms, cc, me = np.empty(self.nnz), np.empty(self.nnz), np.empty(self.nnz)
        self.L.fo_get_csr(self.m, _p(ia), _p(ja), _p(ms), _p(cc), _p(me))
        return dict(ia=ia, ja=ja, mass=ms, curl=cc, mass_eps=me)

    def receivers(self):
        rt = np.empty(self.nR, np.int32)
        n = np.empty((self.nR, 6, 3))
        d = np.empty((self.nR, 6, 3))
        self.L.fo_get_receivers(self.m, _p(rt), _p(n), _p(d))
        return dict(rec_tet=rt, n_arr=n, d_arr=d)

    def source(self):
        se = np.empty(self.nseg, np.int32)
        sd = np.empty(self.nseg, np.int32)
        self.L.fo_get_source(self.m, _p(se), _p(sd))
        return dict(src_edg=se, src_drc=sd)

    def material(self):
        s = np.empty(self.nT)
        im = np.empty(self.nT)
        self.L.fo_get_material(self.m, _p(s), _p(im))
        return s, im

    def tensors(self):
        """(sigma tensors, curl weights) [nT][3][3] of an anisotropic model, or None."""
        s9 = np.empty((self.nT, 3, 3))
        n9 = np.empty((self.nT, 3, 3))
        return (s9, n9) if self.L.fo_get_tensors(self.m, _p(s9), _p(n9)) else None

    def cpu_costs(self, dt, nrhs_block=64):
        """Timed pieces at time step dt: phase 11, phase 22, phase 33 with
        1 and nrhs_block right-hand sides, rhs + SpMV (s), factor nnz, mtype."""
        out = np.zeros(8)
        rc = self.L.fo_cpu_costs(self.m, dt, nrhs_block, _p(out))
        if rc != 0:
            raise RuntimeError(f"PARDISO error {rc}")
        keys = ["phase11_s", "phase22_s", "phase33_s", "phase33_block_s", "rhs_s", "factor_nnz", "mtype"]
        return dict(zip(keys, out[:7].tolist()))

    def set_replay(self, keep):
        k = np.ascontiguousarray(keep, np.int32)
        self.L.fo_set_replay(self.m, len(k), _p(k))

    def set_sigma(self, sigma):
        s = np.ascontiguousarray(sigma, np.float64)
        self.L.fo_set_sigma(self.m, _p(s))

    def sig_parameter(self):
        return self.L.fo_sig_parameter(self.m)

    def set_slices(self, times):
        t = np.ascontiguousarray(times, np.float64)
        self.L.fo_set_slices(self.m, len(t), _p(t))

    def tet_fields(self, e):
        """Per-tet [T][6] E and curl E at the centroids for edge values e[nE]."""
        e = np.ascontiguousarray(e, np.float64)
        out = np.empty((self.nT, 6))
        self.L.fo_tet_fields(self.m, _p(e), _p(out))
        return out

    def slice(self, i):
        """Per-tet fields [T][9] of slice i and the (step, t) of the row that took it."""
        f = np.empty((self.nT, 9))
        st, t = ctypes.c_int(), ctypes.c_double()
        if self.L.fo_get_slice(self.m, i, _p(f), ctypes.byref(st), ctypes.byref(t)) != 0:
            raise RuntimeError(f"slice {i} was not reached")
        return f, st.value, t.value

    def run(self, field=None, field2=None, max_steps=0):
        rc = self.L.fo_run(self.m, field.encode() if field else None, field2.encode() if field2 else None,
                           max_steps)
        if rc != 0:
            raise RuntimeError(self.L.fo_error().decode())
        n = self.L.fo_num_rows(self.m)
        st = np.empty(n, np.int32)
        t, dt = np.empty(n), np.empty(n)
        f = np.empty((n, self.nR, 9))
        self.L.fo_get_rows(self.m, _p(st), _p(t), _p(dt), _p(f))
        nc = self.L.fo_num_checks(self.m)
        cs = np.empty(nc, np.int32)
        ce = np.empty(nc)
        ck = np.empty(nc, np.int32)
        self.L.fo_get_checks(self.m, _p(cs), _p(ce), _p(ck))
        tm = np.empty(6)
        self.L.fo_get_timers(self.m, _p(tm))
        return dict(step=st, t=t, dt=dt, fields=f, chk_step=cs, chk_err=ce, chk_keep=ck, timers=tm,
                    nfact=self.L.fo_num_factorizations(self.m), factor_nnz=self.L.fo_factor_nnz(self.m))


def set_threads(n):
    lib().fo_set_threads(int(n))


def pardiso_solve(ia, ja, va, b):
    """Solve with MKL PARDISO (mtype 2 -> -2), upper CSR 1-based."""
    L = lib()
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    nrhs = 1 if b.ndim == 1 else b.shape[0]
    x = np.empty_like(b)
    rc = L.fo_pardiso_solve_once(len(ia) - 1, _p(ia), _p(ja), _p(va), _p(b), _p(x), nrhs)
    if rc != 0:
        raise RuntimeError(f"PARDISO error {rc}")
    return x


def pardiso_available():
    return bool(lib().fo_pardiso_available())


def tet_matrices(x, y, z, s_mass=1.0, s_curl=1.0):
    F = np.empty(36)
    K = np.empty(36)
    x, y, z = [np.ascontiguousarray(v, np.float64) for v in (x, y, z)]
    lib().fo_tet_matrices(_p(x), _p(y), _p(z), _D(s_mass), _D(s_curl), _p(F), _p(K))
    return F.reshape(6, 6), K.reshape(6, 6)


def tet_cofactors(x, y, z):
    x, y, z = [np.ascontiguousarray(v, np.float64) for v in (x, y, z)]
    a, b, c, d = [np.empty(4) for _ in range(4)]
    vol = _D()
    lib().fo_tet_cofactors(_p(x), _p(y), _p(z), _p(a), _p(b), _p(c), _p(d), ctypes.byref(vol))
    return a, b, c, d, vol.value


def tet_matrices_ani(x, y, z, A_mass, A_curl):
    F = np.empty(36)
    K = np.empty(36)
    x, y, z = [np.ascontiguousarray(v, np.float64) for v in (x, y, z)]
    Am = np.ascontiguousarray(A_mass, np.float64).reshape(9)
    Ak = np.ascontiguousarray(A_curl, np.float64).reshape(9)
    lib().fo_tet_matrices_ani(_p(x), _p(y), _p(z), _p(Am), _p(Ak), _p(F), _p(K))
    return F.reshape(6, 6), K.reshape(6, 6)


def aniso_tensor(values, angles_deg, scale=1.0):
    v = np.ascontiguousarray(values, np.float64)
    a = np.ascontiguousarray(angles_deg, np.float64)
    A = np.empty(9)
    lib().fo_aniso_tensor(_p(v), _p(a), _D(scale), _p(A))
    return A.reshape(3, 3)


def gate_rows(t, f, gates):
    """The .field2 gate rule on given rows: t[nrows], f[nrows][nq] ->
    (values [ngate][nq], filling row per gate or -1)."""
    t = np.ascontiguousarray(t, np.float64)
    f = np.ascontiguousarray(f, np.float64).reshape(len(t), -1)
    g = np.ascontiguousarray(gates, np.float64)
    out = np.full((len(g), f.shape[1]), np.nan)
    row = np.empty(len(g), np.int32)
    lib().fo_gate_rows(len(t), f.shape[1], _p(t), _p(f), len(g), _p(g), _p(out), _p(row))
    return out, row


def source_didt(source_type, pulse_type, t, t0, cur):
    return lib().fo_source_didt(source_type, pulse_type, t, t0, cur)
