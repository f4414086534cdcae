"""pwh_eig_zheevd_noalloc (scratch in a device-resident array, no allocation) against
pwh_eig_zheevd (the same solver on library-allocated scratch) through the C ABI.

Two calls in a row on device memory, the second with a smaller matrix and a padded
leading dimension: each must give the existing entry point's eigenvalues and
eigenvectors (FP64 bar: 1e-12; both run the same kernels on the same input, so they
are expected to agree bitwise).  Sizes above the scratch's limit are refused.  Device
memory for the test's own buffers comes from the HIP runtime the library links."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip_runtime():
    h = C.CDLL(os.path.join("/opt/rocm/lib", "libamdhip64.so"))
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def hermitian(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (A + A.conj().T)


def test_noalloc_eigensolver_matches_allocating_entry():
    import pwh
    pwh.init(0)
    hip = hip_runtime()
    nmax = 128
    dA = DevBuf(hip, 16 * nmax * nmax + 16 * 64 * nmax)
    dW = DevBuf(hip, 8 * nmax)
    dI = DevBuf(hip, 4)
    try:
        for n, ld, seed in ((128, 128, 1), (41, 50, 2)):
            A = hermitian(n, seed)
            w_ref, V_ref = pwh.eig_zheevd(A, ldA=ld)
            buf = np.zeros((n, ld), np.complex128)       # column-major, column j at buf[j, :ld]
            buf[:, :n] = A.T
            dA.put(buf)
            dI.put(np.array([-7], np.int32))
            print(f"no-alloc eigensolve n={n} ldA={ld}", flush=True)
            pwh.eig_zheevd_noalloc(dA.p.value, ld, dW.p.value, n, dI.p.value)
            pwh.synchronize()
            info = int(dI.get(np.int32, 1)[0])
            w = dW.get(np.float64, n)
            V = dA.get(np.complex128, n * ld).reshape(n, ld)[:, :n].T
            assert info == 0
            assert np.abs(w - w_ref).max() < 1e-12 * max(1.0, np.abs(w_ref).max())
            assert np.abs(V - V_ref).max() < 1e-12
            assert np.abs(A @ V - V * w).max() < 1e-12 * np.abs(w).max()
        import pwh as P
        with pytest.raises(P.PwhError) as e:
            pwh.eig_zheevd_noalloc(dA.p.value, 129, dW.p.value, 129, dI.p.value)
        assert e.value.code == P.PWH_ERR_ARG and "exceeds 128" in str(e.value)
    finally:
        dA.free()
        dW.free()
        dI.free()
