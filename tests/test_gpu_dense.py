"""GPU parity of the dense routines next to the hot path that the shim also replaces:
fft_fftw3 (pwh_fft3d) and eig_zheevd (pwh_eig_zheevd), called through the C ABI and
through the reference's mangled C++ symbols.

Tolerances: the FFT within 1e-13 (norm-relative) of numpy's FFT and of the oracle's
mixed-radix DFT, with the fft_fftw3.cpp:27-38 conventions (inverse exp(+i) unscaled,
forward exp(-i) times 1/N, x fastest); eigenvalues within 1e-12 of the reference's own
zheevd output (golden fixture) and of numpy, eigenvectors checked through A V = V Lambda
and V^H V = I to 1e-12 (their phases are a LAPACK implementation choice).
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG, ROOT, relerr

pytestmark = pytest.mark.gpu

SYM_FFT = "_Z9fft_fftw3PCdiiib"
SYM_EIG = "_Z10eig_zheevdPCdiPdi"


def _grid(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("nr", [(12, 16, 20), (32, 32, 32), (24, 30, 36), (64, 64, 64), (96, 96, 96),
                                (128, 20, 12)])
def test_fft3d_matches_numpy(nr):
    import pwh
    pwh.init(0)
    x = _grid((nr[2], nr[1], nr[0]), seed=sum(nr))
    n = x.size
    assert relerr(pwh.fft3d(x, inverse=True), np.fft.ifftn(x) * n) < 1e-13
    y = pwh.fft3d(x, inverse=False)
    assert relerr(y, np.fft.fftn(x) / n) < 1e-13
    assert relerr(pwh.fft3d(y, inverse=True), x) < 1e-13


def test_fft3d_matches_oracle(oracle):
    import pwh
    nr = (20, 12, 16)
    x = _grid((nr[2], nr[1], nr[0]), seed=5)
    for inv in (True, False):
        ref = oracle.fft3d(x.reshape(-1), nr, inv).reshape(x.shape)
        assert relerr(pwh.fft3d(x, inverse=inv), ref) < 1e-13


def test_fft3d_plane_wave_closed_form():
    """exp(+i 2pi (g.r)) on the grid: the forward transform is a unit spike at g."""
    import pwh
    nr = (24, 30, 36)
    g = (5, -7, 11)
    i3, i2, i1 = np.meshgrid(np.arange(nr[2]), np.arange(nr[1]), np.arange(nr[0]), indexing="ij")
    ph = 2 * np.pi * (g[0] * i1 / nr[0] + g[1] * i2 / nr[1] + g[2] * i3 / nr[2])
    y = pwh.fft3d(np.exp(1j * ph), inverse=False)
    spike = np.zeros_like(y)
    spike[g[2] % nr[2], g[1] % nr[1], g[0] % nr[0]] = 1.0
    assert np.abs(y - spike).max() < 1e-13


@pytest.mark.parametrize("nr", [(22, 12, 16), (14, 26, 38), (82, 6, 4), (41, 43, 2), (140, 70, 10),
                                (121, 13, 7), (500, 3, 2)])
def test_fft3d_any_length_matches_numpy(nr):
    """Lengths outside the compiled set (runtime plans: radix 7, and primes 11, 13, 19, 41, 43
    as direct stages)."""
    import pwh
    pwh.init(0)
    x = _grid((nr[2], nr[1], nr[0]), seed=sum(nr) + 1)
    n = x.size
    assert relerr(pwh.fft3d(x, inverse=True), np.fft.ifftn(x) * n) < 1e-13
    assert relerr(pwh.fft3d(x, inverse=False), np.fft.fftn(x) / n) < 1e-13


def test_fft3d_unsupported_length():
    import pwh
    with pytest.raises(pwh.PwhError) as e:
        pwh.fft3d(np.zeros((8, 8, 1), np.complex128), inverse=True)
    assert e.value.code == pwh.PWH_ERR_UNSUPPORTED


def _check_eig(A, lam, V, tol=1e-12):
    s = max(1.0, np.abs(A).max())
    assert np.abs(A @ V - V * lam[None, :]).max() < tol * s * A.shape[0]
    assert np.abs(V.conj().T @ V - np.eye(A.shape[0])).max() < tol * A.shape[0]
    assert np.all(np.diff(lam) >= 0)


def test_eig_zheevd_reference_fixtures():
    import pwh
    pwh.init(0)
    z = np.load(os.path.join(ROOT, "tests", "golden", "ref_lapack.npz"))
    for A, w in ((z["T"], z["w"]), (z["A4"], z["w4"])):
        lam, V = pwh.eig_zheevd(A)
        assert np.abs(lam - w).max() < 1e-12
        H = np.triu(A) + np.triu(A, 1).conj().T
        _check_eig(H, lam, V)


@pytest.mark.parametrize("n,ld", [(128, 128), (97, 110), (432, 432)])
def test_eig_zheevd_random_reads_upper_triangle(n, ld):
    import pwh
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = 0.5 * (B + B.conj().T)
    A = np.triu(H) + np.tril(rng.standard_normal((n, n)), -1)   # garbage below the diagonal
    lam, V = pwh.eig_zheevd(A, ldA=ld)
    assert np.abs(lam - np.linalg.eigvalsh(H)).max() < 1e-12 * np.abs(H).max() * np.sqrt(n)
    _check_eig(H, lam, V, tol=1e-13)


@pytest.fixture(scope="module")
def shim_lib():
    C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    vp, ci = C.c_void_p, C.c_int
    getattr(S, SYM_FFT).argtypes = [vp, ci, ci, ci, C.c_bool]
    getattr(S, SYM_EIG).argtypes = [vp, ci, vp, ci]
    return S


def test_shim_fft_fftw3(shim_lib):
    nr = (32, 24, 20)
    x = _grid((nr[2], nr[1], nr[0]), seed=8)
    d = x.copy()
    getattr(shim_lib, SYM_FFT)(d.ctypes.data, nr[0], nr[1], nr[2], True)
    assert relerr(d, np.fft.ifftn(x) * x.size) < 1e-13
    getattr(shim_lib, SYM_FFT)(d.ctypes.data, nr[0], nr[1], nr[2], False)
    assert relerr(d, x) < 1e-13


def test_shim_eig_zheevd(shim_lib):
    z = np.load(os.path.join(ROOT, "tests", "golden", "ref_lapack.npz"))
    A = z["T"]
    n = A.shape[0]
    buf = np.ascontiguousarray(A.T)          # column-major, ldA = n
    lam = np.zeros(n)
    getattr(shim_lib, SYM_EIG)(buf.ctypes.data, n, lam.ctypes.data, n)
    assert np.abs(lam - z["w"]).max() < 1e-12
    V = buf.T
    _check_eig(np.triu(A) + np.triu(A, 1).conj().T, lam, V)


def _unitary(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return Q


@pytest.mark.parametrize("case", ["4-fold", "40-fold", "near-diagonal", "diagonal", "wide", "n1", "n2",
                                  "n33", "zero"])
def test_eig_zheevd_hard_spectra(case):
    """The one-workgroup solver (n <= 128): exact and near degeneracies (cluster
    Gram-Schmidt), diagonal input (zero off-diagonals), tiny and odd sizes, a zero matrix,
    and a spread like the Rayleigh-Ritz matrices (-0.3 .. 60 Ha)."""
    import pwh
    n = {"n1": 1, "n2": 2, "n33": 33}.get(case, 128)
    Q = _unitary(n, 7)
    rng = np.random.default_rng(3)
    if case == "4-fold":
        lam = np.repeat(np.linspace(-1, 2, 32), 4) + 1e-15 * rng.standard_normal(n)
    elif case == "40-fold":
        lam = np.sort(np.concatenate([np.full(40, 0.3), np.linspace(1, 5, 88)]))
    elif case == "wide":
        lam = np.sort(np.concatenate([np.linspace(-0.3, 0.5, 100), rng.uniform(1, 60, 28)]))
    elif case == "zero":
        lam = np.zeros(n)
    else:
        lam = np.sort(rng.standard_normal(n))
    if case == "diagonal":
        H = np.diag(lam).astype(np.complex128)
    elif case == "near-diagonal":
        H = np.diag(np.linspace(-2, 3, n)) + 1e-9 * (Q + Q.conj().T)
    else:
        H = (Q * lam) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    got, V = pwh.eig_zheevd(H)
    s = max(np.abs(H).max(), 1.0)
    assert np.abs(got - np.linalg.eigvalsh(H)).max() < 1e-12 * s
    _check_eig(H, got, V, tol=1e-13)


@pytest.mark.parametrize("n,case", [(129, "random"), (200, "4-fold"), (256, "diagonal"), (300, "near-diagonal"),
                                    (384, "zero"), (512, "wide"), (640, "random")])
def test_eig_zheevd_mid_sizes(n, case):
    """128 < n: the multi-workgroup Householder reduction, then rocSOLVER's zstedc and
    zunmtr.  Eigenvalues against numpy to 1e-12 of the largest entry, A V = V Lambda and
    V^H V = I to 1e-13 n, on random, degenerate, diagonal, near-diagonal, zero and
    Rayleigh-Ritz-like spectra."""
    import pwh
    Q = _unitary(n, n)
    rng = np.random.default_rng(n + 1)
    if case == "4-fold":
        lam = np.repeat(np.linspace(-1, 2, n // 4), 4)
    elif case == "wide":
        lam = np.sort(np.concatenate([np.linspace(-0.3, 0.5, n - 100), rng.uniform(1, 60, 100)]))
    elif case == "zero":
        lam = np.zeros(n)
    else:
        lam = np.sort(rng.standard_normal(n))
    if case == "diagonal":
        H = np.diag(lam).astype(np.complex128)
    elif case == "near-diagonal":
        H = np.diag(np.linspace(-2, 3, n)) + 1e-9 * (Q + Q.conj().T)
    else:
        H = (Q * lam) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    got, V = pwh.eig_zheevd(H)
    s = max(np.abs(H).max(), 1.0)
    assert np.abs(got - np.linalg.eigvalsh(H)).max() < 1e-12 * s
    _check_eig(H, got, V, tol=1e-13)


def test_shim_eig_zheevd_vectors_only(shim_lib):
    """eig_zheevd.cpp:37-69: the two-argument overload leaves the eigenvectors in A."""
    sym = "_Z10eig_zheevdPCdii"
    getattr(shim_lib, sym).argtypes = [C.c_void_p, C.c_int, C.c_int]
    z = np.load(os.path.join(ROOT, "tests", "golden", "ref_lapack.npz"))
    A = z["T"]
    n = A.shape[0]
    buf = np.ascontiguousarray(A.T)
    getattr(shim_lib, sym)(buf.ctypes.data, n, n)
    _check_eig(np.triu(A) + np.triu(A, 1).conj().T, z["w"], buf.T)
