"""mxm_dg_max_wave_speed_noalloc through the C ABI: the allocation-free
variant of the device CFL reduction (its tile partials in the library's
fixed scratch array, the result in caller-owned device memory) gives the
same bits as mxm_dg_max_wave_speed, on two calls in a row whose second has
the smaller input (fewer tiles and a smaller speed, so partials left over
from the first call would show), in both arithmetic forms and for a 2D and
a 3D solver; a null output is rejected."""
import ctypes as C

import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import _lib
from mojoxm_amd import physics as ph

from test_gpu_parity import random_state

pytestmark = pytest.mark.gpu

HIP_D2H = 2


def hip_runtime():
    """The HIP runtime libmxm_hip.so is linked against (already loaded)."""
    _lib.lib()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def speed_both_ways(s, hip, buf):
    want = s.max_wave_speed()
    s.max_wave_speed_noalloc(buf.value)
    s.synchronize()
    got = C.c_double(-1.0)
    assert hip.hipMemcpy(C.byref(got), buf, 8, HIP_D2H) == 0
    return got.value, want


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_speed_equals_allocating_entry_point(dim, precision):
    hip = hip_runtime()
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), 8) == 0
    try:
        rng = np.random.default_rng(11 + dim)
        if dim == 2:
            sizes, L, P = [(40, 36), (6, 5)], (1.0, 1.0), 3
        else:
            sizes, L, P = [(8, 7, 6), (2, 3, 2)], (1.0, 1.0, 1.0), 2
        results = []
        for n, boost in zip(sizes, (4.0, 1.0)):
            phys = ph.Euler(flux_type=ph.FLUX_HLLEC) if dim == 2 else ph.Euler(flux_type=ph.FLUX_RUSANOV)
            with M.Solver(n, L, phys, P=P, precision=precision) as s:
                q = random_state(phys, s.num_elements, s.NP, rng, dim)
                q[..., 1:4] *= boost  # the first (larger) mesh also has the larger speed
                q[..., 4] = q[..., 4] + 0.5 * (boost ** 2 - 1.0) * (q[..., 1:4] ** 2).sum(-1) / boost ** 2 / q[..., 0]
                if precision == M.PREC_FP32_MFMA:
                    q = q.astype(np.float32).astype(np.float64)
                s.set_state(q)
                got, want = speed_both_ways(s, hip, buf)
                assert np.isfinite(want) and want > 0.0
                assert got == want, (n, got, want)
                results.append(got)
        assert results[1] < results[0]
    finally:
        hip.hipFree(buf)


def test_noalloc_speed_rejects_null_output():
    with M.Solver((4, 4), (1.0, 1.0), ph.Euler(), P=2) as s:
        with pytest.raises(_lib.MxmError):
            s.max_wave_speed_noalloc(0)
