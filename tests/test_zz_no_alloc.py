"""ps3d_diagnose_noalloc (no allocation, static device scratch) against
ps3d_diagnose on the same device fields, through the C-ABI: the same bits,
on two calls in a row, the second with a smaller grid (the scratch keeps the
first call's partial records, which must not leak into the second)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(nx, ny, nz, seed):
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.random_state(nx, ny, nz, seed=seed, lid=True)
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    m.run(2)
    m.synchronize()
    return m


def _noalloc(m):
    from ps3d_amd import _lib
    out = np.zeros(8)
    nan = ctypes.c_int32(-1)
    ptrs = [ctypes.c_void_p(m.device_ptr(n)) for n in ("u", "v", "w", "b")]
    _lib.check(m.lib.ps3d_diagnose_noalloc(ctypes.byref(m._cprm), *ptrs, None,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ctypes.byref(nan)))
    return out, nan.value


@pytest.mark.parametrize("shapes", [((96, 80, 40), (24, 16, 8))])
def test_diagnose_noalloc_matches_diagnose(shapes):
    for q, (nx, ny, nz) in enumerate(shapes):
        print(f"diagnose_noalloc call {q + 1}: {nx}x{ny}x{nz}", flush=True)
        m = _model(nx, ny, nz, seed=11 + q)
        ref = m.diagnose()
        got, nan = _noalloc(m)
        assert nan == 0 and not ref["nan"]
        assert np.array_equal(got, np.array(ref["values"])), (got, ref["values"])
        m.close()


def test_diagnose_noalloc_rejects_bad_input():
    from ps3d_amd import _lib
    from ps3d_amd.params import CParams
    m = _model(16, 16, 8, seed=5)
    out = np.zeros(8)
    nan = ctypes.c_int32()
    dp = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ptrs = [ctypes.c_void_p(m.device_ptr(n)) for n in ("u", "v", "w", "b")]
    bad = CParams.from_params(m.prm.replace(nx=(1 << 20) + 2))
    assert m.lib.ps3d_diagnose_noalloc(ctypes.byref(bad), *ptrs, None, dp, ctypes.byref(nan)) == _lib.PS3D_ERR_ARG
    assert "2^20" in _lib.last_error()
    assert m.lib.ps3d_diagnose_noalloc(ctypes.byref(m._cprm), ptrs[0], None, ptrs[2], ptrs[3], None, dp,
                                       ctypes.byref(nan)) == _lib.PS3D_ERR_ARG
    m.close()
