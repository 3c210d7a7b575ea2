"""tf_solve_noalloc (the allocation-free single right-hand-side solve) through
the C ABI against tf_solve on the same factor: identical results on two
calls in a row, the second on a smaller system, so that the shared scratch
array's leftovers from the first call cannot leak into the second; systems
too large for the scratch array and host pointers are rejected."""
import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu


class DevBuf:
    """Device memory from the HIP runtime the library itself links (the
    test allocates; the entry point under test does not)."""

    def __init__(self, n):
        import ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = n
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n)) == 0

    @property
    def ptr(self):
        return self.p.value

    def put(self, a):
        import ctypes
        a = np.ascontiguousarray(a, np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), 1) == 0

    def get(self):
        import ctypes
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(8 * self.n), 2) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _check(fe, dt, seed):
    fe.factor(0, dt)
    b = np.random.default_rng(seed).standard_normal(fe.nE)
    x_ref = fe.solve(0, b)
    bd, xd = DevBuf(fe.nE), DevBuf(fe.nE)
    bd.put(b)
    xd.put(np.full(fe.nE, np.nan))
    fe.solve_noalloc(0, bd.ptr, xd.ptr)
    x = xd.get()
    bd.free()
    xd.free()
    print(f"[gpu] no-alloc solve n={fe.nE}: max |x - x_ref| {np.abs(x - x_ref).max():.3e}", flush=True)
    np.testing.assert_array_equal(x, x_ref)


def test_noalloc_solve_matches_tf_solve_twice(src_mod_big):
    from temf3dt_amd import fetd, workloads
    fe, _, _ = fetd.setup_from_files(src_mod_big["mesh"], src_mod_big["cntl"])
    _check(fe, 1e-8, 1)
    fe.close()
    small = gpu_from_model(workloads.small_model())  # second call: a smaller system
    assert small.nE < 96712
    _check(small, 1e-7, 2)
    # host pointers are refused with a clear error, not a fault
    b = np.ones(small.nE)
    x = np.empty(small.nE)
    with pytest.raises(fetd.FetdError) as ex:
        small.solve_noalloc(0, b.ctypes.data, x.ctypes.data)
    assert ex.value.code == fetd.TF_ERR_ARG and "device pointers" in str(ex.value)
    small.close()


def test_noalloc_solve_rejects_systems_beyond_the_scratch_array():
    from temf3dt_amd import fetd, workloads
    fe = gpu_from_model(workloads.batched_3d_model(), loops=[0])  # 1.6M unknowns
    fe.factor(0, 1e-8)
    lay = fe.factor_layout()
    need = 4 * fe.nE + lay["struct_rows"]
    assert need * 8 > 64 * 2 ** 20
    bd, xd = DevBuf(fe.nE), DevBuf(fe.nE)
    with pytest.raises(fetd.FetdError) as ex:
        fe.solve_noalloc(0, bd.ptr, xd.ptr)
    assert ex.value.code == fetd.TF_ERR_ARG and "too large" in str(ex.value)
    bd.free()
    xd.free()
    fe.close()
