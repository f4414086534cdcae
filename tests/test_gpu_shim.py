"""The drop-in shim called exactly as the reference's drivers call it: the C++-mangled
ApplyHam_block / ApplyHam / chebyfilt / ortho_qr / lanczos, reading the reference
globals (set here through libpwh_refglobals.so).  Parity with the FP64 oracle: 1e-12
for ApplyHam_block and ApplyHam, 1e-11 for chebyfilt, 1e-10 (absolute, Ha) for the
Lanczos T and 1e-8 for its residual vector."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG, oracle_globals, relerr

pytestmark = pytest.mark.gpu

SYM_BLOCK = "_Z14ApplyHam_blockPCdS0_iiPdii"
SYM_ONE = "_Z8ApplyHamPCdS0_iiPdi"
SYM_CHEB = "_Z9chebyfiltiPCdiddPdi"
SYM_QR = "_Z8ortho_qrPCdii"
SYM_LANCZOS = "_Z7lanczosiPCdiPdiS0_S1_i"


@pytest.fixture(scope="module")
def shim(si8):
    G = C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    g = si8.globals()
    keep = dict(ngw=np.array([si8.ngw], np.int32), n123=np.ascontiguousarray(si8.gv["n123"]),
                xkg=np.ascontiguousarray(si8.gv["xkg"]), na=g.na.copy(), lmax=g.l_max.copy(),
                lloc=g.l_loc.copy(), wnl=g.wnl.copy(), eigr=np.ascontiguousarray(g.eigr),
                pkg=np.ascontiguousarray(g.pkg_a))
    for name, val in (("NR1", si8.nr[0]), ("NR2", si8.nr[1]), ("NR3", si8.nr[2]),
                      ("NNR", int(np.prod(si8.nr))), ("NGWX", g.ngwx), ("NSP", 1), ("NAX", g.nax),
                      ("NLMAX", 4), ("NKPT", 1), ("NX", 14)):
        C.c_int.in_dll(G, name).value = int(val)
    C.c_double.in_dll(G, "TPIBA2").value = float(g.tpiba2)
    for name, arr in (("NGW", "ngw"), ("N123", "n123"), ("XKG", "xkg"), ("NA", "na"),
                      ("L_MAX", "lmax"), ("L_LOC", "lloc"), ("WNL", "wnl"), ("EIGR", "eigr"),
                      ("PKG_A", "pkg")):
        C.c_void_p.in_dll(G, name).value = keep[arr].ctypes.data
    vp, ci, cd = C.c_void_p, C.c_int, C.c_double
    getattr(S, SYM_BLOCK).argtypes = [vp, vp, ci, ci, vp, ci, ci]
    getattr(S, SYM_ONE).argtypes = [vp, vp, ci, ci, vp, ci]
    getattr(S, SYM_CHEB).argtypes = [ci, vp, ci, cd, cd, vp, ci]
    getattr(S, SYM_QR).argtypes = [vp, ci, ci]
    getattr(S, SYM_LANCZOS).argtypes = [ci, vp, ci, vp, ci, vp, vp, ci]
    yield S, keep
    del keep


def test_shim_apply_ham_block(si8, shim, oracle):
    S, _ = shim
    psi = si8.psi(16, seed=31)
    out = np.zeros_like(psi)
    rho = np.ascontiguousarray(si8.vloc)
    getattr(S, SYM_BLOCK)(psi.ctypes.data, out.ctypes.data, si8.ngw, 1, rho.ctypes.data, rho.size, 16)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    assert relerr(out, ref) < 1e-12
    one = np.zeros(si8.ngw, np.complex128)
    v = np.ascontiguousarray(psi[3])
    getattr(S, SYM_ONE)(v.ctypes.data, one.ctypes.data, si8.ngw, 1, rho.ctypes.data, rho.size)
    assert relerr(one, ref[3]) < 1e-12


def test_shim_chebyfilt_uses_nx_plus_2_bands(si8, shim, oracle):
    S, _ = shim
    X = si8.psi(16, seed=32)                 # NX + 2 = 16 (chebyfilt.cpp:20-21)
    ref = oracle.chebyfilt(oracle_globals(si8), X, 5, 0.5, 20.0, si8.vloc, si8.ngw)
    rho = np.ascontiguousarray(si8.vloc)
    getattr(S, SYM_CHEB)(1, X.ctypes.data, 5, 0.5, 20.0, rho.ctypes.data, rho.size)
    assert relerr(X, ref) < 1e-11


def test_shim_ortho_qr(si8, shim):
    S, _ = shim
    rng = np.random.default_rng(33)
    X = rng.standard_normal((16, si8.ngw)) + 1j * rng.standard_normal((16, si8.ngw))
    X0 = X.copy()
    getattr(S, SYM_QR)(X.ctypes.data, si8.ngw, 16)
    assert np.abs(X.conj() @ X.T - np.eye(16)).max() < 1e-13
    Qr, _ = np.linalg.qr(X0.T)
    assert np.abs(X.T @ X.conj() - Qr @ Qr.conj().T).max() < 1e-12


def test_shim_lanczos(si8, shim, oracle):
    """lanczos.cpp:8-101 through its mangled symbol: v0 normalised in place, T[0..k*k)
    cleared then filled (upper + subdiagonal), f the final residual."""
    S, _ = shim
    rng = np.random.default_rng(34)
    v0 = rng.random(si8.ngw) + 1j * rng.random(si8.ngw)
    k = 12
    Tr, fr, _ = oracle.lanczos(oracle_globals(si8), v0, k, si8.vloc, si8.ngw)
    v = np.ascontiguousarray(v0.copy())
    T = np.full(k * k, 7.0)                     # every entry is overwritten or cleared
    f = np.zeros(si8.ngw, np.complex128)
    rho = np.ascontiguousarray(si8.vloc)
    getattr(S, SYM_LANCZOS)(1, v.ctypes.data, k, T.ctypes.data, k, f.ctypes.data, rho.ctypes.data,
                            rho.size)
    assert np.abs(T.reshape(k, k).T - Tr).max() < 1e-10
    assert relerr(f, fr) < 1e-8
    assert abs(np.linalg.norm(v) - 1.0) < 1e-14
