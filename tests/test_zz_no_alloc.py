"""sav_irfft_noalloc (the allocation-free entry point) against sav_irfft through the C ABI.

Device buffers come from the HIP runtime itself (hipMalloc in the test, outside the call
under test).  Two calls in a row, the second with fewer fields, each bit-identical to the
host-buffer entry point (same kernels, same order; only the intermediate's address
differs); the largest count that fits the scratch at 1024^2 works and one more is refused
with SAV_E_ARG and an error text."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    lib = ct.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    lib.hipFree.argtypes = [ct.c_void_p]
    lib.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ct.c_void_p()
        assert hip.hipMalloc(ct.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.asfortranarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.zeros(shape, dtype=dtype, order="F")
        assert out.nbytes <= self.n
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _spectra(conf, count, seed):
    rng = np.random.default_rng(seed)
    shape = (conf.Nxh, conf.Ny) + ((conf.Nz,) if conf.ndim == 3 else ()) + (count,)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _check(jv, hip, conf, counts, precision=0):
    ops = jv.Operators(conf, precision=precision)
    lib = ops.lib
    real = np.float64 if precision != 1 else np.float32
    cplx = np.complex128 if precision != 1 else np.complex64
    P = conf.Nx * conf.Ny * (conf.Nz if conf.ndim == 3 else 1)
    H = conf.Nxh * conf.Ny * (conf.Nz if conf.ndim == 3 else 1)
    cmax = max(counts)
    din = DevBuf(hip, cmax * H * np.dtype(cplx).itemsize)
    dout = DevBuf(hip, cmax * P * np.dtype(real).itemsize)
    try:
        for i, count in enumerate(counts):
            uh = _spectra(conf, count, 100 + i)
            ref = ops.irfft(uh)                      # the host-buffer entry point (sav_irfft)
            din.put(uh.astype(cplx))
            print(f"sav_irfft_noalloc {conf.Nx}^{conf.ndim} count {count} precision {precision}", flush=True)
            rc = lib.sav_irfft_noalloc(ops.h, din.p, dout.p, count)
            assert rc == 0, lib.sav_last_error(ops.h).decode()
            got = dout.get(ref.shape, real)
            if precision == 1:
                # sav_irfft rounds the host input to FP32 the same way: identical kernels
                np.testing.assert_array_equal(got, ref.astype(np.float32))
            else:
                np.testing.assert_array_equal(got, ref)
    finally:
        din.free()
        dout.free()
        ops.close()


def test_noalloc_irfft_equals_irfft_2d(jv, hip):
    _check(jv, hip, jv.workload_c1(n=64, steps=1).conf, [3, 1])


def test_noalloc_irfft_equals_irfft_3d(jv, hip):
    _check(jv, hip, jv.workload_c4(n=16, steps=1).conf, [4, 2])


def test_noalloc_irfft_equals_irfft_fp32(jv, hip):
    _check(jv, hip, jv.workload_c1(n=96, steps=1).conf, [2, 1], precision=1)


def test_noalloc_irfft_scratch_limit_1024(jv, hip):
    conf = jv.workload_c3(n=1024, steps=1).conf
    H = conf.Nxh * conf.Ny
    limit = (64 << 20) // (16 * H)
    assert limit == 7
    _check(jv, hip, conf, [limit, 2])
    ops = jv.Operators(conf)
    d = DevBuf(hip, 256)
    try:
        rc = ops.lib.sav_irfft_noalloc(ops.h, d.p, d.p, limit + 1)
        assert rc == jv.SAV_E_ARG
        assert "scratch" in ops.lib.sav_last_error(ops.h).decode()
        assert ops.lib.sav_irfft_noalloc(ops.h, None, d.p, 1) == jv.SAV_E_ARG
    finally:
        d.free()
        ops.close()
