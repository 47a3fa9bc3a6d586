"""The allocation-free multiply (lfg_spmv_noalloc) against lfg_spmv through the C ABI.

The caller owns the device vectors (allocated here with hipMalloc; the entry point
allocates nothing). Each layout and precision: y bit-identical to lfg_spmv's on the
same x, and x.y equal to the dot product of the two host vectors to 1e-12 relative;
two systems in a row on the same device (the second smaller), so the shared scratch
array is reused with a smaller grid; and the refusal before a layout exists.
"""
import ctypes

import numpy as np
import pytest

from lightfem_amd import meshgen, native, pipeline

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, hip, nbytes: int):
        self.hip, self.ptr = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a: np.ndarray):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0

    def download(self, a: np.ndarray):
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(a.nbytes), D2H) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


def check_system(hip, n_mesh: int, cases):
    dev = native.Device(0)
    try:
        pipeline.setup(meshgen.loguniform_cantilever(n_mesh), device=dev, assembly="coloring")
        n = dev.n
        rng = np.random.default_rng(n_mesh)
        x = rng.standard_normal(n)
        xb, yb, db = DeviceBuffer(hip, 8 * n), DeviceBuffer(hip, 8 * n), DeviceBuffer(hip, 8)
        try:
            xb.upload(x)
            for precision, variant in cases:
                print(f"case noalloc {n_mesh}x{n_mesh} {precision} variant {variant}", flush=True)
                y_ref = dev.spmv(x, precision, variant)  # builds the layout, then the reference result
                dev.spmv_noalloc(xb.ptr.value, yb.ptr.value, db.ptr.value, precision, variant)
                y, xy = np.empty(n), np.empty(1)
                yb.download(y)
                db.download(xy)
                assert np.array_equal(y.view(np.int64), y_ref.view(np.int64)), (precision, variant)
                expect = float(np.dot(x, y_ref))
                assert abs(xy[0] - expect) <= 1e-12 * np.abs(x * y_ref).sum(), (precision, variant)
        finally:
            xb.free(), yb.free(), db.free()
    finally:
        dev.close()


CASES = [("fp64", 4), ("fp64", 6), ("fp64", 8), ("fp64", 9), ("fp64", 15), ("mixed", 4), ("mixed", 6), ("mixed", 8), ("mixed", 15)]


def test_noalloc_matches_spmv_then_smaller_system(hip):
    check_system(hip, 96, CASES)
    check_system(hip, 20, CASES)


def test_noalloc_refuses_without_layout(hip):
    print("case noalloc refusal", flush=True)
    dev = native.Device(0)
    try:
        pipeline.setup(meshgen.loguniform_cantilever(12), device=dev, assembly="coloring")
        xb, yb = DeviceBuffer(hip, 8 * dev.n), DeviceBuffer(hip, 8 * dev.n)
        try:
            rc = dev.spmv_noalloc(xb.ptr.value, yb.ptr.value, 0, "mixed", 6, raise_on_error=False)
            assert rc == native.ERR_STATE
            rc = dev.spmv_noalloc(0, yb.ptr.value, 0, "fp64", 4, raise_on_error=False)
            assert rc == native.ERR_ARG
        finally:
            xb.free(), yb.free()
    finally:
        dev.close()
