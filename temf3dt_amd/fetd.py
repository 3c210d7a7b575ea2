"""Python host side of the FETD forward path over the libtemfetd C ABI.

`Fetd` wraps one tf_ctx (include/temfetd.h).  `run_root` mirrors the
reference driver `FETD -f ROOT` (src/FETD_MAIN.f90:110-165): it reads
ROOT.mesh and ROOT.cntl, assembles, runs the BDF2 time loop on the GPU and
writes ROOT.field / ROOT.field2 in the reference's format.

There is no CPU fallback: if libtemfetd.so is missing or no GPU is present,
the constructor raises.
"""
import ctypes
import os

import numpy as np

from . import io as tfio

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtemfetd.so")
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SIGS = {
    "tf_create": [ctypes.POINTER(_P), _I],
    "tf_destroy": [_P],
    "tf_last_error": [_P],
    "tf_set_mesh": [_P, _I, _P, _I, _P],
    "tf_get_sizes": [_P, _P, _P],
    "tf_get_topology": [_P, _P, _P, _P, _P, _P],
    "tf_set_materials": [_P, _P, _P],
    "tf_set_source": [_P, _I, _I, _P, _I, _I, _D, _D],
    "tf_set_receivers": [_P, _I, _P, _D, _P],
    "tf_set_schedule": [_P, _I, _I, _I, _D],
    "tf_assemble": [_P],
    "tf_get_matrices": [_P, _P, _P],
    "tf_get_receiver_vectors": [_P, _P, _P],
    "tf_analyze": [_P, _I],
    "tf_factor_stats": [_P, _P, _P, _P, _P],
    "tf_factor_pivots": [_P, _P, _P, _P],
    "tf_factor_layout": [_P, _P, _P, _P],
    "tf_solve_noalloc": [_P, _I, _P, _P],
    "tf_factor": [_P, _I, _D],
    "tf_factor_csr": [_P, _I, _I, _P, _P, _P],
    "tf_solve": [_P, _I, _I, _P, _P],
    "tf_release": [_P, _I],
    "tf_run": [_P, _I, _I, _P, _P, _P, _P, _P],
    "tf_get_checks": [_P, _I, _P, _P, _P, _P],
    "tf_get_timers": [_P, _P],
    "tf_set_num_sources": [_P, _I],
    "tf_set_replay": [_P, _I, _P],
    "tf_set_slices": [_P, _I, _P],
    "tf_get_slice": [_P, _I, _P, _P, _P],
    "tf_set_tuning": [_P, ctypes.c_char_p, _I],
    "tf_time_solve": [_P, _I, _I, _I, _P],
    "tf_set_profiling": [_P, _I],
    "tf_get_profile": [_P, _I, _P, _P],
    "tf_set_checkpointing": [_P, _I],
    "tf_adjoint": [_P, _I, _P, _P],
    "tf_set_materials_tensor": [_P, _P, _P],
    "tf_set_models": [_P, _I, _P, _P],
    "tf_run_models": [_P, _I, _I],
    "tf_get_model_rows": [_P, _I, _I, _P, _P, _P, _P, _P],
}

# kernel-timing categories of tf_get_profile (include/temfetd.h)
PROF_CATEGORIES = ["tile", "tile_red", "pull", "factor_update", "factor_other", "solve1", "step"]


def load_library(path=LIB_PATH):
    """Load libtemfetd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m temf3dt_amd.build`")
    lib = ctypes.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _I
    lib.tf_destroy.restype = None
    lib.tf_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class FetdError(RuntimeError):
    """A failed tf_* call; `code` is its status (TF_ERR_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


# status codes (include/temfetd.h)
TF_ERR_ARG, TF_ERR_STATE, TF_ERR_HIP, TF_ERR_NUMERIC, TF_ERR_INPUT = -1, -2, -3, -4, -5


class Fetd:
    """One FETD model on one GPU (a tf_ctx)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.ctx = _P()
        rc = self.lib.tf_create(ctypes.byref(self.ctx), device)
        if rc != 0:
            msg = self.lib.tf_last_error(None).decode(errors="replace")
            raise FetdError(f"tf_create failed ({rc}) for GPU device {device}: {msg}")
        self.nR = 0
        self.nsrc = 1

    def close(self):
        if self.ctx:
            self.lib.tf_destroy(self.ctx)
            self.ctx = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.tf_last_error(self.ctx).decode(errors="replace")
            raise FetdError(f"{what} failed ({rc}): {msg}", rc)

    # -- model_mesh -------------------------------------------------------
    def set_mesh(self, xyz, tet_vex):
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tet_vex = np.ascontiguousarray(tet_vex, np.int32)
        self._chk(self.lib.tf_set_mesh(self.ctx, len(self.xyz), _ptr(self.xyz), len(self.tet_vex),
                                       _ptr(self.tet_vex)), "tf_set_mesh")
        ne, nz = _I(), _I()
        self._chk(self.lib.tf_get_sizes(self.ctx, ctypes.byref(ne), ctypes.byref(nz)), "tf_get_sizes")
        self.nE, self.nnz = ne.value, nz.value

    def topology(self):
        T = len(self.tet_vex)
        edg = np.empty((self.nE, 2), np.int32)
        te = np.empty((T, 6), np.int32)
        td = np.empty((T, 6), np.int32)
        ia = np.empty(self.nE + 1, np.int32)
        ja = np.empty(self.nnz, np.int32)
        self._chk(self.lib.tf_get_topology(self.ctx, _ptr(edg), _ptr(te), _ptr(td), _ptr(ia), _ptr(ja)),
                  "tf_get_topology")
        return dict(edg_vex=edg, tet_edg=te, tet_drc=td, ia=ia, ja=ja)

    # -- model_ctl --------------------------------------------------------
    def set_materials(self, sigma, inv_mu):
        s = np.ascontiguousarray(sigma, np.float64)
        m = np.ascontiguousarray(inv_mu, np.float64)
        self._chk(self.lib.tf_set_materials(self.ctx, _ptr(s), _ptr(m)), "tf_set_materials")

    def set_materials_tensor(self, sig9, nu9):
        """Anisotropic media: [nT][3][3] conductivity tensors and curl weights."""
        self.sig9 = np.ascontiguousarray(sig9, np.float64).reshape(-1, 9)
        self.nu9 = np.ascontiguousarray(nu9, np.float64).reshape(-1, 9)
        self._chk(self.lib.tf_set_materials_tensor(self.ctx, _ptr(self.sig9), _ptr(self.nu9)),
                  "tf_set_materials_tensor")

    def set_source(self, seg_vex, source_type, pulse_type, pulse_width, current, isrc=0):
        sv = np.ascontiguousarray(seg_vex, np.int32).reshape(-1, 2)
        self._chk(self.lib.tf_set_source(self.ctx, isrc, len(sv), _ptr(sv), source_type, pulse_type,
                                         pulse_width, current), "tf_set_source")

    def set_sources(self, loops, source_type, pulse_type, pulse_width, current):
        """Several transmitter loops sharing one waveform (block solve)."""
        self._chk(self.lib.tf_set_num_sources(self.ctx, len(loops)), "tf_set_num_sources")
        self.nsrc = len(loops)
        for q, segs in enumerate(loops):
            self.set_source(segs, source_type, pulse_type, pulse_width, current, isrc=q)

    def set_replay(self, keep):
        k = np.ascontiguousarray(keep if keep is not None else [], np.int32)
        self._chk(self.lib.tf_set_replay(self.ctx, len(k), _ptr(k) if len(k) else None), "tf_set_replay")

    def set_slices(self, times):
        """Per-tet field slices at these times (the .cntl vtkout list)."""
        t = np.ascontiguousarray(times if times is not None else [], np.float64)
        self.nslice = len(t)
        self._chk(self.lib.tf_set_slices(self.ctx, len(t), _ptr(t) if len(t) else None), "tf_set_slices")

    def slice(self, i):
        """Slice i after a run: per-tet [nT][9] (E, curl E, B at the centroid)
        and the (step, t) of the output row that took it."""
        f = np.empty((len(self.tet_vex), 9))
        st, t = _I(), ctypes.c_double()
        self._chk(self.lib.tf_get_slice(self.ctx, i, _ptr(f), ctypes.byref(st), ctypes.byref(t)), "tf_get_slice")
        return f, st.value, t.value

    def set_receivers(self, xyz, rec_charac=10.0):
        r = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        rt = np.empty(len(r), np.int32)
        self._chk(self.lib.tf_set_receivers(self.ctx, len(r), _ptr(r), rec_charac, _ptr(rt)), "tf_set_receivers")
        self.nR = len(r)
        return rt

    def set_schedule(self, sp_division, sp_double, sp_dblesize, time_maximum):
        self._chk(self.lib.tf_set_schedule(self.ctx, sp_division, sp_double, sp_dblesize, time_maximum),
                  "tf_set_schedule")

    # -- mesh_process -------------------------------------------------------
    def assemble(self):
        self._chk(self.lib.tf_assemble(self.ctx), "tf_assemble")

    def matrices(self):
        M = np.empty(self.nnz)
        K = np.empty(self.nnz)
        self._chk(self.lib.tf_get_matrices(self.ctx, _ptr(M), _ptr(K)), "tf_get_matrices")
        return M, K

    def receiver_vectors(self):
        n = np.empty((self.nR, 6, 3))
        d = np.empty((self.nR, 6, 3))
        self._chk(self.lib.tf_get_receiver_vectors(self.ctx, _ptr(n), _ptr(d)), "tf_get_receiver_vectors")
        return n, d

    # -- solver seam ----------------------------------------------------------
    def analyze(self, leaf_size=0):
        self._chk(self.lib.tf_analyze(self.ctx, leaf_size), "tf_analyze")

    def set_tuning(self, key, value):
        """Block-solve tile-kernel knob (see tf_set_tuning)."""
        self._chk(self.lib.tf_set_tuning(self.ctx, key.encode(), int(value)), f"tf_set_tuning({key})")

    def time_solve(self, slot, nrhs, reps=3):
        """Average ms of a device-resident solve with nrhs right-hand sides."""
        ms = _D()
        self._chk(self.lib.tf_time_solve(self.ctx, slot, nrhs, reps, ctypes.byref(ms)), "tf_time_solve")
        return ms.value

    def set_profiling(self, on=True):
        self._chk(self.lib.tf_set_profiling(self.ctx, 1 if on else 0), "tf_set_profiling")

    def profile(self):
        """{category: (launches, ms)} of the kernels timed since set_profiling."""
        out = {}
        for k, name in enumerate(PROF_CATEGORIES):
            n, ms = ctypes.c_int64(), _D()
            self._chk(self.lib.tf_get_profile(self.ctx, k, ctypes.byref(n), ctypes.byref(ms)), "tf_get_profile")
            out[name] = (n.value, ms.value)
        return out

    def factor_stats(self):
        nnzL = ctypes.c_int64()
        fl = _D()
        ns, nl = _I(), _I()
        self._chk(self.lib.tf_factor_stats(self.ctx, ctypes.byref(nnzL), ctypes.byref(fl), ctypes.byref(ns),
                                           ctypes.byref(nl)), "tf_factor_stats")
        return dict(nnzL=nnzL.value, flops=fl.value, nsuper=ns.value, nlevels=nl.value)

    def factor_layout(self):
        """Struct row indices, L doubles per slot, update-stack doubles."""
        v = [ctypes.c_int64() for _ in range(3)]
        self._chk(self.lib.tf_factor_layout(self.ctx, *[ctypes.byref(x) for x in v]), "tf_factor_layout")
        return dict(struct_rows=v[0].value, l_doubles=v[1].value, u_doubles=v[2].value)

    def factor_pivots(self):
        """Pivots of the last factorisation: negative count, min and max |d|."""
        nn = ctypes.c_int64()
        lo, hi = _D(), _D()
        self._chk(self.lib.tf_factor_pivots(self.ctx, ctypes.byref(nn), ctypes.byref(lo), ctypes.byref(hi)),
                  "tf_factor_pivots")
        return dict(negative=nn.value, min_abs=lo.value, max_abs=hi.value)

    def factor(self, slot, dt):
        self._chk(self.lib.tf_factor(self.ctx, slot, dt), "tf_factor")

    def factor_csr(self, slot, ia, ja, va):
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        va = np.ascontiguousarray(va, np.float64)
        self._chk(self.lib.tf_factor_csr(self.ctx, slot, len(ia) - 1, _ptr(ia), _ptr(ja), _ptr(va)),
                  "tf_factor_csr")

    def solve(self, slot, b):
        b = np.ascontiguousarray(b, np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        x = np.empty_like(b)
        self._chk(self.lib.tf_solve(self.ctx, slot, nrhs, _ptr(b), _ptr(x)), "tf_solve")
        return x

    def solve_noalloc(self, slot, b_ptr, x_ptr):
        """tf_solve_noalloc: x = P^-1 b with b and x device addresses (ints,
        n doubles each, edge order); allocates nothing."""
        self._chk(self.lib.tf_solve_noalloc(self.ctx, slot, _P(b_ptr), _P(x_ptr)), "tf_solve_noalloc")

    def release(self, slot):
        self._chk(self.lib.tf_release(self.ctx, slot), "tf_release")

    # -- time_solver ---------------------------------------------------------
    def run(self, max_steps=0, max_rows=20000):
        n = _I()
        steps = np.empty(max_rows, np.int32)
        t = np.empty(max_rows)
        dt = np.empty(max_rows)
        f = np.empty((max_rows, self.nsrc, self.nR, 9))
        self._chk(self.lib.tf_run(self.ctx, max_steps, max_rows, ctypes.byref(n), _ptr(steps), _ptr(t), _ptr(dt),
                                  _ptr(f)), "tf_run")
        k = min(n.value, max_rows)
        fields = f[:k].copy()
        return dict(step=steps[:k].copy(), t=t[:k].copy(), dt=dt[:k].copy(),
                    fields=fields[:, 0] if self.nsrc == 1 else fields)

    # -- independent models -------------------------------------------------------
    def set_models(self, ntet, nrec):
        ntet = np.ascontiguousarray(ntet, np.int32)
        nrec = np.ascontiguousarray(nrec, np.int32)
        self._chk(self.lib.tf_set_models(self.ctx, len(ntet), _ptr(ntet), _ptr(nrec)), "tf_set_models")
        self.model_nrec = nrec.copy()

    def run_models(self, max_steps=0, graphs=True):
        self._chk(self.lib.tf_run_models(self.ctx, max_steps, 1 if graphs else 0), "tf_run_models")

    def model_rows(self, m, max_rows=20000):
        nr = int(self.model_nrec[m])
        n = _I()
        steps = np.empty(max_rows, np.int32)
        t = np.empty(max_rows)
        dt = np.empty(max_rows)
        f = np.empty((max_rows, nr, 9))
        self._chk(self.lib.tf_get_model_rows(self.ctx, m, max_rows, ctypes.byref(n), _ptr(steps), _ptr(t), _ptr(dt),
                                             _ptr(f)), "tf_get_model_rows")
        k = min(n.value, max_rows)
        return dict(step=steps[:k].copy(), t=t[:k].copy(), dt=dt[:k].copy(), fields=f[:k].copy())

    # -- adjoint ----------------------------------------------------------------
    def set_checkpointing(self, every):
        """Record the next run's tape with history checkpoints every `every`
        steps (0: off), for adjoint()."""
        self._chk(self.lib.tf_set_checkpointing(self.ctx, int(every)), "tf_set_checkpointing")

    def adjoint(self, w):
        """dJ/dlog(sigma) per tet for J = sum(w * fields) over the rows of the
        last (taped) run; w shaped like run()'s fields."""
        w = np.ascontiguousarray(w, np.float64)
        nrows = w.shape[0]
        g = np.empty(len(self.tet_vex))
        self._chk(self.lib.tf_adjoint(self.ctx, nrows, _ptr(w), _ptr(g)), "tf_adjoint")
        return g

    def checks(self):
        n = _I()
        self._chk(self.lib.tf_get_checks(self.ctx, 0, ctypes.byref(n), None, None, None), "tf_get_checks")
        st = np.empty(n.value, np.int32)
        er = np.empty(n.value)
        kp = np.empty(n.value, np.int32)
        self._chk(self.lib.tf_get_checks(self.ctx, n.value, ctypes.byref(n), _ptr(st), _ptr(er), _ptr(kp)),
                  "tf_get_checks")
        return dict(step=st, err=er, keep=kp)

    def timers(self):
        t = np.empty(8)
        self._chk(self.lib.tf_get_timers(self.ctx, _ptr(t)), "tf_get_timers")
        keys = ["assemble", "factor", "solve", "rhs", "readout", "loop", "analyze", "nfactor"]
        return dict(zip(keys, t.tolist()))


def setup_from_files(mesh_path, cntl_path, device=0, sig_rescale=False):
    """model_mesh + model_ctl + mesh_process from reference input files.
    sig_rescale: the small-loop variant (src_small/FETD_FWD.f90:314-328):
    sigma of domains 2..n scaled by sp = 1 / sigma of the last domain (c["sp"])."""
    m = tfio.read_mesh(mesh_path)
    c = tfio.read_cntl(cntl_path)
    if c["sp_mthd"] != 2:
        raise FetdError("error!")  # only the adaptive BDF2 method exists (src/FETD_FWD.f90:1210-1219)
    if c["bound_cond"] != 0:
        raise FetdError("absorbing boundary condition is not supported")
    c["sp"] = None
    fe = Fetd(device)
    fe.set_mesh(m["xyz"], m["tet_vex"])
    if c.get("aniso"):
        if sig_rescale:
            raise FetdError("the conductivity rescaling applies to isotropic domain tables")
        fe.set_materials_tensor(*tfio.materials_tensor(m["tet_phy"], c["aniso"]))
    else:
        if sig_rescale:
            c["sp"] = tfio.sig_parameter(c["domains"])
        fe.set_materials(*tfio.materials(m["tet_phy"], c["domains"], sig_rescale))
    fe.set_source(m["seg_vex"], c["source_type"], c["pulse_type"], c["pulse_width"], c["current"])
    fe.set_receivers(c["receivers"], c["rec_charac"])
    fe.set_schedule(c["sp_division"], c["sp_double"], c["sp_dblesize"], c["time_maximum"])
    if c["vtk_times"]:
        fe.set_slices(c["vtk_times"])
    fe.assemble()
    return fe, m, c


def run_root(root, device=0, write=True, max_steps=0, sig_rescale=False, replay=None):
    """`FETD -f ROOT`: returns the rows and writes ROOT.field / ROOT.field2,
    and ROOT_i.msh for every vtkout slice the run reached when ROOT.msh exists
    (msh_output, src/FETD_FWD.f90:1642-1656,1713-1864).  replay: keep/change
    outcomes of the doubling checks to follow instead of computing them."""
    fe, m, c = setup_from_files(root + ".mesh", root + ".cntl", device, sig_rescale)
    if replay is not None:
        fe.set_replay(replay)
    rows = fe.run(max_steps=max_steps)
    if write:
        sp = c["sp"]
        tfio.write_field(root + ".field", rows["step"], rows["t"], rows["fields"], c["output"], sp)
        if c["time_points"]:
            tfio.write_field2(root + ".field2", rows["t"], rows["fields"], c["time_points"], c["output"], sp)
        if c["vtk_times"] and os.path.exists(root + ".msh"):
            for i in range(len(c["vtk_times"])):
                if rows["t"][-1] < c["vtk_times"][i]:
                    continue  # not reached
                f9, _, _ = fe.slice(i)
                tfio.write_msh_slice(root + ".msh", tfio.slice_path(root, i + 1), f9, c["output"])
    rows["timers"] = fe.timers()
    fe.close()
    return rows
