"""e2d_run_loop_noalloc (the run loop on caller-owned arrays, with its scratch
in a static device array) against e2d_run_loop on a solver handle.

The caller's arrays are allocated here with hipMalloc through ctypes (the
entry point itself allocates nothing) in the library's device layout, filled
with the handle's initial state, advanced by both entry points and compared
bit for bit (ghost cells included), with t and nStep.  Two calls in a row on
the same GPU, the second on a smaller grid, then a split-precision call.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import (E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64, E2D_U, E2D_U2,
                                E2DError, HydroParams, HydroRun)
from euler2d_kokkos_amd._capi import E2D_LAYOUT_LEFT, check, load_library

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DeviceState:
    """One state array of the caller, in the library's device layout:
    element (v, j, i) at v * plane + j * pitch + i."""

    def __init__(self, hip, p, precision):
        self.hip, self.p, self.precision = hip, p, precision
        self.pitch = (p.isize + 15) // 16 * 16
        self.plane = (self.pitch * p.jsize + 63) // 64 * 64
        self.shape = (4, self.plane, 2) if precision == E2D_PRECISION_FP32_SPLIT else (4, self.plane)
        self.dtype = np.float32 if precision == E2D_PRECISION_FP32_SPLIT else np.float64
        self.nbytes = int(np.prod(self.shape)) * np.dtype(self.dtype).itemsize
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        zero = np.zeros(self.shape, self.dtype)
        assert hip.hipMemcpy(self.ptr, zero.ctypes.data, self.nbytes, H2D) == 0

    def upload(self, x):
        """x: (4, jsize, isize) float64 values (split: hi = float(x),
        lo = float(x - hi), as the library's own upload)."""
        host = np.zeros(self.shape, self.dtype)
        p = self.p
        for v in range(4):
            for j in range(p.jsize):
                row = x[v, j]
                o = j * self.pitch
                if self.precision == E2D_PRECISION_FP32_SPLIT:
                    hi = row.astype(np.float32)
                    lo = (row - hi.astype(np.float64)).astype(np.float32)
                    host[v, o:o + p.isize, 0] = hi
                    host[v, o:o + p.isize, 1] = lo
                else:
                    host[v, o:o + p.isize] = row
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def download(self):
        host = np.empty(self.shape, self.dtype)
        assert self.hip.hipMemcpy(host.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        p = self.p
        out = np.empty((4, p.jsize, p.isize))
        for v in range(4):
            for j in range(p.jsize):
                o = j * self.pitch
                if self.precision == E2D_PRECISION_FP32_SPLIT:
                    out[v, j] = (host[v, o:o + p.isize, 0].astype(np.float64)
                                 + host[v, o:o + p.isize, 1].astype(np.float64))
                else:
                    out[v, j] = host[v, o:o + p.isize]
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(p, precision, a, b, steps, t=0.0, n=0, seg_rows=0):
    lib = load_library()
    tt, nn = C.c_double(t), C.c_int(n)
    rc = lib.e2d_run_loop_noalloc(C.byref(p.raw), precision, a.ptr, b.ptr, a.pitch, a.plane,
                                  seg_rows, steps, C.byref(tt), C.byref(nn), None)
    check(rc)
    return tt.value, nn.value


def _compare(params, precision, steps, seg_rows, tag):
    hip = _hip()
    p = HydroParams(params.copy())
    ref = HydroRun(HydroParams(params.copy()), precision=precision)
    x0 = ref.get_state(E2D_U, E2D_LAYOUT_LEFT)
    if precision == E2D_PRECISION_FP32_SPLIT:
        # both runs start from the same doubles, split by the same rule
        ref.set_state(E2D_U, x0, E2D_LAYOUT_LEFT)
        ref.set_state(E2D_U2, x0, E2D_LAYOUT_LEFT)
    ref.run(steps)
    a, b = DeviceState(hip, p, precision), DeviceState(hip, p, precision)
    try:
        a.upload(x0)
        b.upload(x0)
        t, n = _noalloc(p, precision, a, b, steps, seg_rows=seg_rows)
        print(f"{tag}: noalloc t={t!r} nStep={n}; handle t={ref.t!r} nStep={ref.nStep}",
              flush=True)
        assert (t, n) == (ref.t, ref.nStep) and n == steps
        np.testing.assert_array_equal(a.download(), ref.get_state(E2D_U, E2D_LAYOUT_LEFT),
                                      err_msg=tag)
        np.testing.assert_array_equal(b.download(), ref.get_state(E2D_U2, E2D_LAYOUT_LEFT),
                                      err_msg=tag)
    finally:
        a.free()
        b.free()
        ref.close()


def test_noalloc_equals_run_loop(gpu_available):
    # first call: a blast on 200x150 (four strips, several segments)
    p1 = O.default_params(nx=200, ny=150, problemType=1, cfl=0.7, nStepmax=10**6, tEnd=1e9)
    _compare(p1, E2D_PRECISION_FP64, 40, 0, "blast 200x150 fp64")
    # second call on the same GPU's scratch, smaller grid, other BCs and tile
    p2 = O.default_params(nx=40, ny=36, problemType=0, cfl=0.8, nStepmax=10**6, tEnd=1e9,
                          boundary_type_xmin=2, boundary_type_xmax=2)
    _compare(p2, E2D_PRECISION_FP64, 30, 7, "implode 40x36 fp64")
    # split precision
    p3 = O.default_params(nx=96, ny=64, problemType=1, cfl=0.7, nStepmax=10**6, tEnd=1e9)
    _compare(p3, E2D_PRECISION_FP32_SPLIT, 25, 0, "blast 96x64 split")


def test_noalloc_loop_end_and_errors(gpu_available):
    """tEnd ends the loop inside the budget (clamped last step), as in
    e2d_run_loop; bad layouts are refused before any GPU work."""
    hip = _hip()
    params = O.default_params(nx=48, ny=40, problemType=0, cfl=0.8, nStepmax=10**6, tEnd=0.02)
    p = HydroParams(params.copy())
    ref = HydroRun(HydroParams(params.copy()))
    x0 = ref.get_state(E2D_U, E2D_LAYOUT_LEFT)
    ref.run(0)
    a, b = DeviceState(hip, p, E2D_PRECISION_FP64), DeviceState(hip, p, E2D_PRECISION_FP64)
    try:
        a.upload(x0)
        b.upload(x0)
        t, n = _noalloc(p, E2D_PRECISION_FP64, a, b, 0)
        assert (t, n) == (ref.t, ref.nStep) and t == 0.02
        final = a.download() if n % 2 == 0 else b.download()
        np.testing.assert_array_equal(final, ref.get_state(ref.current(), E2D_LAYOUT_LEFT))
        lib = load_library()
        tt, nn = C.c_double(0.0), C.c_int(0)
        with pytest.raises(E2DError):
            check(lib.e2d_run_loop_noalloc(C.byref(p.raw), E2D_PRECISION_FP64, a.ptr, b.ptr,
                                           p.isize - 1, a.plane, 0, 5, C.byref(tt),
                                           C.byref(nn), None))
        with pytest.raises(E2DError):
            check(lib.e2d_run_loop_noalloc(C.byref(p.raw), E2D_PRECISION_FP64, a.ptr, a.ptr,
                                           a.pitch, a.plane, 0, 5, C.byref(tt), C.byref(nn),
                                           None))
    finally:
        a.free()
        b.free()
        ref.close()
