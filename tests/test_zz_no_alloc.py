"""ifos_opt_pcg_noalloc (include/ifos2d_opt.h), the PCG direction whose
reduction scratch is a static device array, against ifos_opt_pcg with its
allocated context: the same directions bit for bit, the same beta and the
same updated p_last / c_last, on two calls in a row (the second on a smaller
grid, so it runs over the first call's leftover partial sums), for an
iteration that conjugates with p_last (2) and one with c_last (3)."""
import ctypes as C

import numpy as np
import pytest

import ifos_amd
from ifos_amd.opt import Optimizer

pytestmark = pytest.mark.gpu

DB = ifos_amd.DeviceBuffer
_VP = C.c_void_p


def _noalloc(L):
    f = L.ifos_opt_pcg_noalloc
    f.argtypes = [C.c_long, _VP, C.c_float, _VP, _VP, C.c_int, C.c_int, _VP, C.POINTER(C.c_double)]
    f.restype = C.c_int
    return f


def _dev(a):
    return DB.from_array(np.ascontiguousarray(a, np.float32))


def test_noalloc_pcg_equals_context_pcg(gpu_lib):
    f = _noalloc(gpu_lib)
    opt = Optimizer()
    rng = np.random.default_rng(11)
    for (ny, nx), it in (((613, 917), 2), ((97, 131), 3)):
        n = ny * nx
        pl = rng.standard_normal((ny, nx)).astype(np.float32)
        cl = rng.standard_normal((ny, nx)).astype(np.float32)
        g = (rng.standard_normal((ny, nx)) + 0.6 * pl).astype(np.float32)
        # context entry point
        dg, dpl, dcl, dd = _dev(g), _dev(pl), _dev(cl), DB(4 * n)
        beta_ctx = opt.pcg(n, dg, 2.0, dpl, dcl, it, 0, dd)
        want = (dd.to_array((ny, nx)), dpl.to_array((ny, nx)), dcl.to_array((ny, nx)))
        # no-alloc entry point on fresh copies of the same inputs
        dg2, dpl2, dcl2, dd2 = _dev(g), _dev(pl), _dev(cl), DB(4 * n)
        beta = C.c_double(-1.0)
        rc = f(n, dg2.ptr, C.c_float(2.0), dpl2.ptr, dcl2.ptr, it, 0, dd2.ptr, C.byref(beta))
        assert rc == 0, gpu_lib.ifos_last_error().decode()
        assert beta.value == beta_ctx and beta_ctx > 0.0
        np.testing.assert_array_equal(dd2.to_array((ny, nx)), want[0])
        np.testing.assert_array_equal(dpl2.to_array((ny, nx)), want[1])
        np.testing.assert_array_equal(dcl2.to_array((ny, nx)), want[2])
    opt.close()


def test_noalloc_pcg_rejects_oversized_grid(gpu_lib):
    f = _noalloc(gpu_lib)
    beta = C.c_double()
    # rejected on the host before any launch: no device memory is touched
    assert f(1 << 31, None, C.c_float(1.0), None, None, 3, 0, None, C.byref(beta)) == -2
    assert "limit" in gpu_lib.ifos_last_error().decode()
