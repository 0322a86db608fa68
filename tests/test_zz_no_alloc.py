"""The allocation-free calcDtNew entry point against the workspace one.

pennant_hip_calc_dt_new_noalloc keeps its block partials in a library-owned
device array; pennant_hip_calc_dt_new takes a caller workspace.  Both run the
same kernels, so their results must be bit-identical.  Two calls in a row,
the second on fewer zones (a smaller grid over the same array), and a third
large one again, each checked against the workspace entry point and the
closed-form minimum.
"""
import ctypes

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd

pytestmark = pytest.mark.gpu


def _dt_inputs(rng, n):
    zdl = rng.random(n) + 0.5
    zdu = rng.random(n)
    zss = rng.random(n) * 1e-3
    return zdl, zdu, zss


def test_noalloc_dt_new_matches_workspace_entry_point():
    from hip_rt import DeviceArray, synchronize
    L = pa.lib()
    rng = np.random.default_rng(2024)
    with pa.Problem(wd.sedov_rect(16, cstop=1)) as pb:
        base = pb.mesh_desc().contents
        ws = DeviceArray.zeros(L.pennant_hip_reduce_workspace_bytes() // 8)
        # large (more zones than the 1024-block grid covers in one pass), then
        # small, then large again: the second call reuses the array with a
        # smaller grid, the third with a larger one
        for n in (600_000, 1_000, 300_001):
            zdl, zdu, zss = _dt_inputs(rng, n)
            dl, du, ds = DeviceArray(zdl), DeviceArray(zdu), DeviceArray(zss)
            m = pa.Mesh.from_buffer_copy(base)
            m.numz = n
            out_ws, out_na = DeviceArray.zeros(1), DeviceArray.zeros(1)
            synchronize()
            assert L.pennant_hip_calc_dt_new(ctypes.byref(m), dl.ptr, du.ptr, ds.ptr, 0.6, out_ws.ptr,
                                             ws.ptr, None) == 0
            assert L.pennant_hip_calc_dt_new_noalloc(ctypes.byref(m), dl.ptr, du.ptr, ds.ptr, 0.6,
                                                     out_na.ptr, None) == 0
            synchronize()
            a, b = out_ws.numpy()[0], out_na.numpy()[0]
            want = np.min(zdl * 0.6 / np.maximum(zdu, np.maximum(zss, 1e-99)))
            print(f"n={n}: workspace {a!r} noalloc {b!r} numpy {want!r}", flush=True)
            assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)
            assert b == want


def test_noalloc_rejects_bad_arguments():
    from hip_rt import DeviceArray
    L = pa.lib()
    with pa.Problem(wd.sedov_rect(8, cstop=1)) as pb:
        m = pa.Mesh.from_buffer_copy(pb.mesh_desc().contents)
        m.numz = -1
        out = DeviceArray.zeros(1)
        assert L.pennant_hip_calc_dt_new_noalloc(ctypes.byref(m), None, None, None, 0.6,
                                                 out.ptr, None) == pa.EINVAL
        m.numz = 4
        assert L.pennant_hip_calc_dt_new_noalloc(ctypes.byref(m), None, None, None, 0.6, None,
                                                 None) == pa.EINVAL
