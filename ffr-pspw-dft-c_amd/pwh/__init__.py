"""pwh -- Python binding of libpwh, the gfx950 Kohn-Sham H|psi> / ChFSI library.

A thin ctypes layer over the C ABI declared in include/pwh.h.  The names follow the
reference routines they replace (ApplyHam_block, chebyfilt, ortho_qr, lanczos, the
Rayleigh-Ritz and residual blocks of ChebySCFLoop).  Wave-function blocks are numpy
complex128 arrays of shape (nstates, ngw), C-contiguous, which is exactly the
reference's column-major [ngw x nstates] buffer.

There is no fallback: every compute call goes through libpwh.so on the GPU, and
importing this module fails if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libpwh.so")
SHIM_PATH = os.path.join(PKG_DIR, "libpwh_shim.so")
REFGLOBALS_PATH = os.path.join(PKG_DIR, "libpwh_refglobals.so")

FP64, FP32, BF16X3, BF16X6 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "bf16x6": BF16X6}

PWH_OK = 0
PWH_ERR_ARG = -1
PWH_ERR_HIP = -2
PWH_ERR_UNSUPPORTED = -3
PWH_ERR_LAPACK = -4

# Every function declared in include/pwh.h (checked by the CPU test suite).
EXPORTS = [
    "pwh_init", "pwh_last_error", "pwh_device_count", "pwh_selftest_mfma", "pwh_grid_supported",
    "pwh_fft_grid", "pwh_gvec_sizes", "pwh_gvectors", "pwh_phase_factor",
    "pwh_system_create", "pwh_system_destroy", "pwh_system_info", "pwh_set_vloc",
    "pwh_apply_ham_block", "pwh_chebyfilt", "pwh_ortho_qr", "pwh_lanczos", "pwh_rayleigh_ritz",
    "pwh_residual_norms", "pwh_chfsi", "pwh_resident_set", "pwh_resident_get",
    "pwh_resident_apply", "pwh_autotune", "pwh_resident_chebyfilt", "pwh_synchronize", "pwh_timing_enable",
    "pwh_timing_reset", "pwh_timing_get", "pwh_batch_create", "pwh_batch_destroy",
    "pwh_batch_set", "pwh_batch_get", "pwh_batch_chebyfilt", "pwh_fft3d", "pwh_eig_zheevd",
    "pwh_eval_rho_psi", "pwh_davidson", "pwh_lobpcg", "pwh_grid_create", "pwh_grid_destroy",
    "pwh_eval_energy", "pwh_xc_pz", "pwh_normalize_rho", "pwh_mix_simple", "pwh_mix_kerker",
    "pwh_kerker_prec", "pwh_mix_anderson", "pwh_mix_pulay", "pwh_mix_broyden", "pwh_eval_enl",
    "pwh_phase_factor_gpu", "pwh_structure_factor", "pwh_nl_form_factor", "pwh_form_factor",
    "pwh_resident_chfsi", "pwh_resident_eval_rho_psi", "pwh_resident_eval_enl",
    "pwh_eig_zheevd_noalloc", "pwh_fft_path", "pwh_projector_mode", "pwh_tuning_get", "pwh_tuning_set",
    "pwh_scf_state_create", "pwh_scf_state_destroy", "pwh_scf_state_set_rho", "pwh_scf_state_get",
    "pwh_scf_state_density", "pwh_scf_state_mix", "pwh_scf_state_energy", "pwh_scf_state_set_vloc",
]


class PwhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpwh error {code}: {msg}")
        self.code = code


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class SystemDesc(C.Structure):
    _fields_ = [("nr1", C.c_int), ("nr2", C.c_int), ("nr3", C.c_int), ("ngw", C.c_int),
                ("ngwx", C.c_int), ("n123", C.c_void_p), ("xkg", C.c_void_p),
                ("tpiba2", C.c_double), ("nsp", C.c_int), ("nax", C.c_int), ("nlmax", C.c_int),
                ("na", C.c_void_p), ("l_max", C.c_void_p), ("l_loc", C.c_void_p),
                ("wnl", C.c_void_p), ("eigr", C.c_void_p), ("pkg_a", C.c_void_p)]


DIAG_REPORT = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double)
PWH_DIAG_ITER, PWH_DIAG_SMALL_OVERLAP, PWH_DIAG_RANK, PWH_DIAG_EIG_FAILED = 1, 2, 3, 4


class DiagOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("nempty", C.c_int), ("report", DIAG_REPORT),
                ("report_ctx", C.c_void_p), ("iterations", C.c_int), ("nconv", C.c_int),
                ("reason", C.c_int), ("t_total_s", C.c_double), ("band_applies", C.c_longlong)]


class GridDesc(C.Structure):
    _fields_ = [("nr1", C.c_int), ("nr2", C.c_int), ("nr3", C.c_int), ("ng", C.c_int),
                ("n1", C.c_void_p), ("n2", C.c_void_p), ("n3", C.c_void_p), ("g2", C.c_void_p),
                ("nsp", C.c_int), ("sfac", C.c_void_p), ("vps", C.c_void_p), ("rhops", C.c_void_p),
                ("tpiba2", C.c_double), ("omega", C.c_double), ("ngw", C.c_int), ("n123", C.c_void_p),
                ("xkg", C.c_void_p)]


class PseudoDesc(C.Structure):
    _fields_ = [("nsp", C.c_int), ("maxr", C.c_int), ("nang", C.c_int), ("mmax", C.c_void_p),
                ("clog", C.c_void_p), ("r", C.c_void_p), ("psi", C.c_void_p), ("vion", C.c_void_p),
                ("zv", C.c_void_p), ("rgauss", C.c_void_p), ("l_max", C.c_void_p), ("l_loc", C.c_void_p)]


class Energy(C.Structure):
    _fields_ = [("ekin", C.c_double), ("enl", C.c_double), ("esr", C.c_double), ("eself", C.c_double),
                ("etot", C.c_double), ("eht", C.c_double), ("epseu", C.c_double), ("exc", C.c_double),
                ("eh_electron", C.c_double), ("eh_pseudo", C.c_double), ("n_negative", C.c_int)]


class ChfsiOpts(C.Structure):
    _fields_ = [("degree", C.c_int), ("niter", C.c_int), ("precision", C.c_int),
                ("lanczos_k", C.c_int), ("bounds_mode", C.c_int), ("lb", C.c_double),
                ("ub", C.c_double), ("t_filter_s", C.c_double), ("t_ortho_s", C.c_double),
                ("t_rr_s", C.c_double), ("t_lanczos_s", C.c_double), ("t_total_s", C.c_double),
                ("band_applies", C.c_longlong)]


_lib = None


def lib():
    """Load libpwh.so (raises if it has not been built: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          f"(make -C {PKG_DIR})")
    L = C.CDLL(LIB_PATH)
    L.pwh_last_error.restype = C.c_char_p
    vp = C.c_void_p
    proto = {
        "pwh_init": [C.c_int],
        "pwh_device_count": [C.POINTER(C.c_int)],
        "pwh_selftest_mfma": [C.c_int, C.POINTER(C.c_double)],
        "pwh_grid_supported": [C.c_int, C.c_int, C.c_int],
        "pwh_fft_grid": [_dp, C.c_double, _ip],
        "pwh_gvec_sizes": [_dp, C.c_double, _ip, _dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "pwh_gvectors": [_dp, C.c_double, _ip, _dp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                         _ip, _ip, _ip, _dp, _up, _dp],
        "pwh_phase_factor": [_dp, _ip, C.c_int, _ip, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip,
                             C.c_int, _ip, vp],
        "pwh_system_create": [C.POINTER(SystemDesc), C.POINTER(vp)],
        "pwh_system_destroy": [vp],
        "pwh_system_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_size_t)],
        "pwh_set_vloc": [vp, _dp, C.c_int],
        "pwh_fft_path": [vp, C.c_int, C.POINTER(C.c_int)],
        "pwh_apply_ham_block": [vp, vp, vp, C.c_int, C.c_int, C.c_int],
        "pwh_chebyfilt": [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int],
        "pwh_ortho_qr": [vp, C.c_int, C.c_int],
        "pwh_lanczos": [vp, vp, C.c_int, _dp, C.c_int, vp],
        "pwh_rayleigh_ritz": [vp, vp, C.c_int, _dp],
        "pwh_residual_norms": [vp, vp, C.c_int, _dp],
        "pwh_chfsi": [vp, vp, C.c_int, vp, C.POINTER(ChfsiOpts), _dp, _dp],
        "pwh_resident_chfsi": [vp, C.c_int, vp, C.POINTER(ChfsiOpts), _dp, _dp],
        "pwh_resident_eval_rho_psi": [vp, C.c_int, _dp, C.c_double, _dp, C.POINTER(C.c_double), _dp],
        "pwh_resident_eval_enl": [vp, C.c_int, _dp, C.c_double, vp, C.POINTER(C.c_double)],
        "pwh_resident_set": [vp, vp, C.c_int],
        "pwh_resident_get": [vp, vp, C.c_int],
        "pwh_resident_apply": [vp, C.c_int, C.c_int],
        "pwh_autotune": [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)],
        "pwh_projector_mode": [vp, C.c_int, C.c_void_p, C.c_void_p],
        "pwh_tuning_get": [vp, C.c_int, C.c_void_p],
        "pwh_tuning_set": [vp, C.c_int, C.c_void_p],
        "pwh_resident_chebyfilt": [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int],
        "pwh_synchronize": [],
        "pwh_timing_enable": [C.c_int],
        "pwh_timing_reset": [],
        "pwh_timing_get": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)],
        "pwh_batch_create": [C.POINTER(vp), _ip, C.c_int, C.POINTER(vp)],
        "pwh_batch_destroy": [vp],
        "pwh_batch_set": [vp, C.c_int, vp],
        "pwh_batch_get": [vp, C.c_int, vp],
        "pwh_batch_chebyfilt": [vp, C.c_int, _dp, _dp, C.c_int, C.c_int],
        "pwh_fft3d": [vp, C.c_int, C.c_int, C.c_int, C.c_int],
        "pwh_eig_zheevd": [vp, C.c_int, _dp, C.c_int],
        "pwh_eig_zheevd_noalloc": [vp, C.c_int, vp, C.c_int, vp, vp],
        "pwh_eval_rho_psi": [vp, vp, C.c_int, C.c_int, _dp, C.c_double, _dp, C.POINTER(C.c_double), _dp],
        "pwh_davidson": [vp, vp, C.c_int, _dp, C.POINTER(DiagOpts), _dp, _dp],
        "pwh_grid_create": [C.POINTER(GridDesc), C.POINTER(vp)],
        "pwh_grid_destroy": [vp],
        "pwh_eval_energy": [vp, _dp, C.POINTER(Energy)],
        "pwh_xc_pz": [_dp, _dp, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)],
        "pwh_normalize_rho": [_dp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)],
        "pwh_mix_simple": [C.c_double, _dp, _dp, C.c_int, C.POINTER(C.c_double)],
        "pwh_mix_kerker": [vp, C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_double)],
        "pwh_kerker_prec": [vp, _dp, C.c_int, C.c_double, C.c_double],
        "pwh_mix_anderson": [C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_double)],
        "pwh_mix_pulay": [vp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                          C.POINTER(C.c_double)],
        "pwh_mix_broyden": [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                            C.POINTER(C.c_double)],
        "pwh_eval_enl": [vp, vp, C.c_int, C.c_int, _dp, C.c_double, vp, C.POINTER(C.c_double)],
        "pwh_phase_factor_gpu": [_dp, _ip, C.c_int, _ip, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, C.c_int, _ip, vp],
        "pwh_structure_factor": [_dp, _ip, C.c_int, _ip, C.c_int, _dp, C.c_int, C.c_int, _ip, vp],
        "pwh_nl_form_factor": [C.POINTER(PseudoDesc), C.c_double, C.c_double, C.c_int, _ip, C.c_int, _dp, _dp, _dp,
                               C.c_int, _dp, vp, _dp],
        "pwh_form_factor": [C.POINTER(PseudoDesc), _ip, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp,
                            C.POINTER(C.c_double)],
        "pwh_lobpcg": [vp, vp, C.c_int, _dp, C.POINTER(DiagOpts), _dp, _dp],
        "pwh_scf_state_create": [vp, C.c_int, C.c_int, C.POINTER(vp)],
        "pwh_scf_state_destroy": [vp],
        "pwh_scf_state_set_rho": [vp, _dp],
        "pwh_scf_state_get": [vp, vp, vp],
        "pwh_scf_state_density": [vp, vp, C.c_int, _dp, C.c_double, C.POINTER(C.c_double), _dp],
        "pwh_scf_state_mix": [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)],
        "pwh_scf_state_energy": [vp, C.POINTER(Energy)],
        "pwh_scf_state_set_vloc": [vp, vp],
    }
    for name, args in proto.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != PWH_OK:
        raise PwhError(rc, lib().pwh_last_error().decode())


def _zptr(a):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _block(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a if a.ndim == 2 else a.reshape(1, -1)


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


# --------------------------------------------------------------------------
# library / setup
# --------------------------------------------------------------------------

def init(device=0):
    _check(lib().pwh_init(device))


def device_count():
    n = C.c_int()
    rc = lib().pwh_device_count(C.byref(n))
    return n.value if rc == PWH_OK else 0


def selftest_mfma(kind):
    err = C.c_double()
    _check(lib().pwh_selftest_mfma(kind, C.byref(err)))
    return err.value


def grid_supported(nr):
    return bool(lib().pwh_grid_supported(int(nr[0]), int(nr[1]), int(nr[2])))


def _lat(A):
    return np.ascontiguousarray(np.asarray(A, np.float64).reshape(9))


def fft_grid(A, ecut):
    """Setup.cpp:20-34."""
    nr = np.zeros(3, np.int32)
    _check(lib().pwh_fft_grid(_lat(A), float(ecut), nr))
    return tuple(int(x) for x in nr)


def gvectors(A, ecut, nr, xk=None):
    """GVectorsGen0/GVectorsGen/GkVectorsGen: the reference's G maps as its globals hold them."""
    xk = np.zeros((1, 3)) if xk is None else np.asarray(xk, np.float64).reshape(-1, 3)
    nkpt = xk.shape[0]
    xkf = np.ascontiguousarray(xk.reshape(-1))
    nr = np.asarray(nr, np.int32)
    L = lib()
    ngx, ngwx = C.c_int(), C.c_int()
    _check(L.pwh_gvec_sizes(_lat(A), float(ecut), nr, xkf, nkpt, C.byref(ngx), C.byref(ngwx)))
    ngx, ngwx = ngx.value, ngwx.value
    ng = C.c_int()
    in123 = np.zeros(3 * ngx, np.int32)
    ngw = np.zeros(nkpt, np.int32)
    igk = np.zeros(ngwx * nkpt, np.int32)
    xkg = np.zeros(ngwx * nkpt)
    n123 = np.zeros(ngwx * nkpt, np.uint32)
    ggk = np.zeros(3 * ngwx * nkpt)
    _check(L.pwh_gvectors(_lat(A), float(ecut), nr, xkf, nkpt, ngx, ngwx, C.byref(ng), in123, ngw,
                          igk, xkg, n123, ggk))
    return dict(ngx=ngx, ngwx=ngwx, ng=ng.value, in123=in123, ngw=ngw, igk=igk, xkg=xkg,
                n123=n123, ggk=ggk, nkpt=nkpt)


def phase_factor(A, nr, na, nax, tau, gv):
    """PhaseFactor.cpp:5-80 -> EIGR [ngwx*nax*nsp*nkpt] complex128."""
    na = np.asarray(na, np.int32)
    eigr = np.zeros(gv["ngwx"] * nax * len(na) * gv["nkpt"], np.complex128)
    _check(lib().pwh_phase_factor(_lat(A), np.asarray(nr, np.int32), len(na), na, nax,
                                  np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1)),
                                  gv["nkpt"], gv["ngw"], gv["ngwx"], gv["igk"], gv["ngx"],
                                  gv["in123"], _zptr(eigr)))
    return eigr


def phase_factor_gpu(A, nr, na, nax, tau, gv):
    """PhaseFactor.cpp:5-80 on the GPU -> EIGR [ngwx*nax*nsp*nkpt] complex128."""
    na = np.asarray(na, np.int32)
    eigr = np.zeros(gv["ngwx"] * nax * len(na) * gv["nkpt"], np.complex128)
    _check(lib().pwh_phase_factor_gpu(_lat(A), np.asarray(nr, np.int32), len(na), na, nax,
                                      np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1)),
                                      gv["nkpt"], gv["ngw"], gv["ngwx"], gv["igk"], gv["ngx"], gv["in123"],
                                      _zptr(eigr)))
    return eigr


def structure_factor(A, nr, na, nax, tau, gv):
    """StructureFactor.cpp:5-29 on the GPU for the density G-vectors of gv -> [nsp][ng]."""
    na = np.asarray(na, np.int32)
    sf = np.zeros(len(na) * gv["ng"], np.complex128)
    _check(lib().pwh_structure_factor(_lat(A), np.asarray(nr, np.int32), len(na), na, nax,
                                      np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1)), gv["ng"],
                                      gv["ngx"], gv["in123"], _zptr(sf)))
    return sf.reshape(gv["ng"], len(na)).T.copy()


def _pseudo(ps, maxr=570, maxam=3):
    """ctypes description + kept arrays of an oracle-style Pseudo (mmax, clog, r [nsp][m],
    psi, vion [nsp][nang][m], zv, rgauss, l_max, l_loc)."""
    nsp = len(ps.mmax)
    r = np.zeros(maxr * nsp)
    psi = np.zeros(maxr * nsp * maxam)
    vion = np.zeros(maxr * nsp * maxam)
    for i in range(nsp):
        m = int(ps.mmax[i])
        r[i * maxr:i * maxr + m] = ps.r[i][:m]
        for l in range(min(maxam, np.asarray(ps.psi).shape[1])):
            o = (l * nsp + i) * maxr
            psi[o:o + m] = ps.psi[i][l][:m]
            vion[o:o + m] = ps.vion[i][l][:m]
    keep = dict(mmax=np.asarray(ps.mmax, np.int32), clog=np.asarray(ps.clog, np.float64), r=r, psi=psi,
                vion=vion, zv=np.asarray(ps.zv, np.float64), rgauss=np.asarray(ps.rgauss, np.float64),
                l_max=np.asarray(ps.l_max, np.int32), l_loc=np.asarray(ps.l_loc, np.int32))
    d = PseudoDesc(nsp=nsp, maxr=maxr, nang=maxam, **{k: v.ctypes.data for k, v in keep.items()})
    return d, keep


def nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, xk, nlmax=4):
    """NLFormFactor.cpp:5-133 on the GPU for one k-point: (WNL [nsp][nlmax], PKG_A [nsp][ngw][nlmax])."""
    d, keep = _pseudo(ps)
    nsp = len(ps.mmax)
    wnl = np.zeros(nsp * nlmax)
    pkg_a = np.zeros(nlmax * ngw * nsp)
    _check(lib().pwh_nl_form_factor(C.byref(d), float(omega), float(tpiba), 1, np.array([ngw], np.int32), ngw,
                                    np.ascontiguousarray(xkg[:ngw], np.float64),
                                    np.ascontiguousarray(np.asarray(ggk).reshape(-1)[:3 * ngw], np.float64),
                                    np.ascontiguousarray(xk, np.float64), nlmax, wnl, None, pkg_a))
    del keep
    return wnl.reshape(nlmax, nsp).T.copy(), pkg_a.reshape(nsp, ngw, nlmax)


def form_factor(ps, na, omega, tpiba2, g2):
    """FormFactor.cpp:5-72 on the GPU: (VPS [nsp][ng], RHOPS [nsp][ng], ESELF)."""
    d, keep = _pseudo(ps)
    nsp = len(ps.mmax)
    ng = len(g2)
    vps = np.zeros(nsp * ng)
    rhops = np.zeros(nsp * ng)
    es = C.c_double()
    _check(lib().pwh_form_factor(C.byref(d), np.asarray(na, np.int32), float(omega), float(tpiba2), ng,
                                 np.ascontiguousarray(g2, np.float64), vps, rhops, C.byref(es)))
    del keep
    return vps.reshape(ng, nsp).T.copy(), rhops.reshape(ng, nsp).T.copy(), es.value


def ortho_qr(X):
    """ortho_qr.cpp:3-34 on X (nstates, nbasis); returns the orthonormalised copy."""
    X = np.array(_block(X), order="C")
    _check(lib().pwh_ortho_qr(_zptr(X), X.shape[1], X.shape[0]))
    return X


def fft3d(data, inverse):
    """fft_fftw3.cpp:13-46 on a complex grid shaped (nr3, nr2, nr1) (x fastest); returns a copy.

    inverse: exp(+i), unscaled; forward: exp(-i) times 1/(nr1 nr2 nr3)."""
    d = np.array(data, dtype=np.complex128, order="C")
    assert d.ndim == 3
    nr3, nr2, nr1 = d.shape
    _check(lib().pwh_fft3d(_zptr(d), nr1, nr2, nr3, 1 if inverse else 0))
    return d


def eig_zheevd(A, ldA=None):
    """eig_zheevd.cpp:5-34 ('V','U') on the column-major n x n matrix A (numpy (n, n), entry
    [i, j] = A(i, j)); only the upper triangle is read.  Returns (lambda, eigenvectors)."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    ld = n if ldA is None else ldA
    buf = np.zeros((n, ld), np.complex128)      # column j at buf[j, :ld]
    buf[:, :n] = A.T
    lam = np.zeros(n)
    _check(lib().pwh_eig_zheevd(_zptr(buf), ld, lam, n))
    return lam, np.ascontiguousarray(buf[:, :n].T)


def eig_zheevd_noalloc(A_device, ldA, lambda_device, n, info_device, stream=None):
    """pwh_eig_zheevd_noalloc on device addresses (ints); enqueued on `stream` (None: the
    library stream) without waiting."""
    _check(lib().pwh_eig_zheevd_noalloc(C.c_void_p(A_device), int(ldA), C.c_void_p(lambda_device), int(n),
                                        C.c_void_p(info_device), C.c_void_p(stream) if stream else None))


def synchronize():
    _check(lib().pwh_synchronize())


def timing_enable(on=True):
    _check(lib().pwh_timing_enable(1 if on else 0))


def timing_reset():
    _check(lib().pwh_timing_reset())


def timing_get(name):
    ms, n = C.c_double(), C.c_int()
    _check(lib().pwh_timing_get(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


# --------------------------------------------------------------------------
# the device-resident system
# --------------------------------------------------------------------------

@dataclass
class Globals:
    """The reference globals read by ApplyHam_block for one k-point."""
    nr: tuple
    ngw: int
    ngwx: int
    n123: np.ndarray    # uint32 [ngw] (or [ngwx]) 1-based
    xkg: np.ndarray     # float64 [ngw]
    tpiba2: float
    na: np.ndarray      # int32 [nsp]
    nax: int
    l_max: np.ndarray   # int32 [nsp]
    l_loc: np.ndarray   # int32 [nsp]
    nlmax: int
    wnl: np.ndarray     # float64 [nsp*nlmax]
    eigr: np.ndarray    # complex128 [ngwx*nax*nsp]
    pkg_a: np.ndarray   # float64 [nlmax*ngwx*nsp]


class System:
    """A pwh_system: maps, phases and projectors resident in HBM for one k-point."""

    def __init__(self, g: Globals):
        self._keep = [np.ascontiguousarray(g.n123, dtype=np.uint32),
                      np.ascontiguousarray(g.xkg, dtype=np.float64),
                      np.ascontiguousarray(g.na, dtype=np.int32),
                      np.ascontiguousarray(g.l_max, dtype=np.int32),
                      np.ascontiguousarray(g.l_loc, dtype=np.int32),
                      np.ascontiguousarray(g.wnl, dtype=np.float64),
                      np.ascontiguousarray(g.eigr, dtype=np.complex128),
                      np.ascontiguousarray(g.pkg_a, dtype=np.float64)]
        n123, xkg, na, lmx, llc, wnl, eigr, pkg = self._keep
        d = SystemDesc(int(g.nr[0]), int(g.nr[1]), int(g.nr[2]), int(g.ngw), int(g.ngwx),
                       n123.ctypes.data, xkg.ctypes.data, float(g.tpiba2), int(len(na)),
                       int(g.nax), int(g.nlmax), na.ctypes.data, lmx.ctypes.data,
                       llc.ctypes.data, wnl.ctypes.data, eigr.ctypes.data, pkg.ctypes.data)
        h = C.c_void_p()
        _check(lib().pwh_system_create(C.byref(d), C.byref(h)))
        self.h = h
        self.ngw = int(g.ngw)
        self.nnr = int(g.nr[0] * g.nr[1] * g.nr[2])
        # projectors: NA(is) x lm_end(is) with lm_end = L_MAX^2 - 2 L_LOC + 1 (EvalENL.cpp:25)
        lm_end = np.maximum(lmx.astype(np.int64) ** 2 - 2 * llc.astype(np.int64) + 1, 0)
        self.nproj = int(np.sum(na.astype(np.int64) * lm_end))
        self._keep = None

    def close(self):
        if getattr(self, "h", None):
            lib().pwh_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        ngw, ncol, nxa, nbytes = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        _check(lib().pwh_system_info(self.h, C.byref(ngw), C.byref(ncol), C.byref(nxa),
                                     C.byref(nbytes)))
        return dict(ngw=ngw.value, ncol=ncol.value, nxa=nxa.value, device_bytes=nbytes.value)

    FFT_PATHS = {"auto": 0, "compiled": 1, "runtime": 2, "split": 3}

    def fft_path(self, path="auto"):
        """Selects the FFT implementation (pwh_fft_path); returns the mode in effect."""
        used = C.c_int(0)
        _check(lib().pwh_fft_path(self.h, self.FFT_PATHS[path] if isinstance(path, str) else int(path),
                                  C.byref(used)))
        return {v: k for k, v in self.FFT_PATHS.items()}[used.value]

    def set_vloc(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        _check(lib().pwh_set_vloc(self.h, rho, len(rho)))

    def apply_ham_block(self, vin, precision=FP64):
        """ApplyHam_block.cpp:5-172: returns H vin for vin (nstates, ngw)."""
        vin = _block(vin)
        vout = np.empty_like(vin)
        _check(lib().pwh_apply_ham_block(self.h, _zptr(vin), _zptr(vout), vin.shape[1],
                                         vin.shape[0], _prec(precision)))
        return vout

    def chebyfilt(self, X, degree, lb, ub, precision=FP64):
        """chebyfilt.cpp:6-74 on a copy of X (nstates, ngw)."""
        X = np.array(_block(X), order="C")
        _check(lib().pwh_chebyfilt(self.h, _zptr(X), X.shape[0], int(degree), float(lb), float(ub),
                                   _prec(precision)))
        return X

    def eval_rho_psi(self, eigvec, w, omega, rhoe=None, ld=None):
        """EvalRhoPsi.cpp:83-157 on eigvec (nstates, >= ngw) rows = states; w per state.
        Returns (rhoe + density [nnr], EKIN increment, charge per state)."""
        ev = np.ascontiguousarray(eigvec, dtype=np.complex128)
        nst, ldv = ev.shape
        rho = np.zeros(self.nnr) if rhoe is None else np.array(rhoe, dtype=np.float64)
        charge = np.zeros(nst)
        ek = C.c_double(0.0)
        _check(lib().pwh_eval_rho_psi(self.h, _zptr(ev), ldv if ld is None else ld, nst,
                                      np.ascontiguousarray(w, dtype=np.float64), float(omega), rho,
                                      C.byref(ek), charge))
        return rho, ek.value, charge

    def _diag(self, fn, X0, prec, max_iter, nempty):
        X = np.array(_block(X0), order="C")
        n = X.shape[0]
        ev = np.zeros(n)
        res = np.zeros(n)
        events = []
        cb = DIAG_REPORT(lambda ctx, e, a, b, v: events.append((e, a, b, v)))
        o = DiagOpts(max_iter=int(max_iter if max_iter else max(30, n // 2)), nempty=int(nempty), report=cb)
        _check(fn(self.h, _zptr(X), n, np.ascontiguousarray(prec, dtype=np.float64), C.byref(o), ev, res))
        stats = {f: getattr(o, f) for f, _ in DiagOpts._fields_ if f not in ("report", "report_ctx")}
        stats["events"] = events
        return ev, X, res, stats

    def davidson(self, X0, prec, max_iter=None, nempty=0):
        """davidson_cpu.cpp:9-247 from start vectors X0 (nstates, ngw).  Returns (Ritz values,
        final block, residual norms, stats incl. the reported events)."""
        return self._diag(lib().pwh_davidson, X0, prec, max_iter, nempty)

    def lobpcg(self, X0, prec, max_iter=None, nempty=0):
        """lobpcg_cpu.cpp:32-389 from start vectors X0 (nstates, ngw)."""
        return self._diag(lib().pwh_lobpcg, X0, prec, max_iter, nempty)

    def eval_enl(self, c0, focc, wkpt, ld=None):
        """EvalENL.cpp:5-49: returns (ENL, fnl [nproj][nstates] in species, atom, lm row order)."""
        c0 = np.ascontiguousarray(c0, dtype=np.complex128)
        nst, ldc = c0.shape
        enl = C.c_double()
        fnl = np.zeros((max(self.nproj, 1), nst), np.complex128)
        _check(lib().pwh_eval_enl(self.h, _zptr(c0), ldc if ld is None else ld, nst,
                                  np.ascontiguousarray(focc, dtype=np.float64), float(wkpt),
                                  _zptr(fnl), C.byref(enl)))
        return enl.value, fnl[:self.nproj]

    def lanczos(self, v0, k):
        """lanczos.cpp:8-101.  Returns (T[k,k] as the reference fills it, f, v0 normalised)."""
        v0 = np.array(np.asarray(v0, np.complex128).reshape(-1), order="C")
        T = np.zeros(k * k)
        f = np.zeros(self.ngw, np.complex128)
        _check(lib().pwh_lanczos(self.h, _zptr(v0), int(k), T, int(k), _zptr(f)))
        return T.reshape(k, k).T.copy(), f, v0

    def rayleigh_ritz(self, X):
        """ChebySCFLoop.cpp:292-311: returns (eigenvalues, rotated X)."""
        X = np.array(_block(X), order="C")
        w = np.zeros(X.shape[0])
        _check(lib().pwh_rayleigh_ritz(self.h, _zptr(X), X.shape[0], w))
        return w, X

    def residual_norms(self, X):
        """ChebySCFLoop.cpp:415-433."""
        X = _block(X)
        r = np.zeros(X.shape[0])
        _check(lib().pwh_residual_norms(self.h, _zptr(X), X.shape[0], r))
        return r

    def chfsi(self, X, degree=10, niter=10, precision=FP64, lanczos_k=20, bounds_mode=0,
              v0=None, lb=0.0, ub=0.0):
        """ChFSI outer loop (eigen part of ChebySCFLoop.cpp:271-311)."""
        X = np.array(_block(X), order="C")
        if v0 is None:
            v0 = np.ones(self.ngw, np.complex128)
        v0 = np.array(np.asarray(v0, np.complex128).reshape(-1), order="C")
        o = ChfsiOpts(int(degree), int(niter), _prec(precision), int(lanczos_k), int(bounds_mode),
                      float(lb), float(ub), 0, 0, 0, 0, 0, 0)
        eig = np.zeros(X.shape[0])
        res = np.zeros(X.shape[0])
        _check(lib().pwh_chfsi(self.h, _zptr(X), X.shape[0], _zptr(v0), C.byref(o), eig, res))
        stats = {f: getattr(o, f) for f, _ in ChfsiOpts._fields_}
        return eig, X, res, stats

    # device-resident entry points
    def resident_set(self, X):
        X = _block(X)
        _check(lib().pwh_resident_set(self.h, _zptr(X), X.shape[0]))

    def resident_get(self, nstates):
        X = np.zeros((nstates, self.ngw), np.complex128)
        _check(lib().pwh_resident_get(self.h, _zptr(X), nstates))
        return X

    def resident_apply(self, nstates, precision=FP64):
        _check(lib().pwh_resident_apply(self.h, int(nstates), _prec(precision)))

    def projector_mode(self, mode, A=None, tau=None):
        """pwh_projector_mode: 0 stored projector rows, 1 generated on the fly in the FP32
        GEMMs from the factorised phases (A: lattice, tau: cartesian positions)."""
        if mode == 0:
            _check(lib().pwh_projector_mode(self.h, 0, None, None))
        else:
            a = _lat(A)
            tp = np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1))
            _check(lib().pwh_projector_mode(self.h, int(mode), a.ctypes.data, tp.ctypes.data))

    def tuning(self, precision=FP64):
        """pwh_tuning_get: the settings in use for a precision."""
        out = np.zeros(5, np.int32)
        _check(lib().pwh_tuning_get(self.h, _prec(precision), out.ctypes.data))
        return {"target_blocks": int(out[0]), "proj_stage_depth": int(out[1]), "expand_stage_depth": int(out[2]),
                "fft_plan_variant": int(out[3]), "plane_blocks_per_cu": int(out[4])}

    def set_tuning(self, precision, settings):
        """pwh_tuning_set: sets the settings tuning() reports (a dict with its keys; missing
        keys keep the value in use), without timing."""
        cur = self.tuning(precision)
        cur.update(settings)
        keys = ("target_blocks", "proj_stage_depth", "expand_stage_depth", "fft_plan_variant",
                "plane_blocks_per_cu")
        a = np.array([cur[k] for k in keys], np.int32)
        _check(lib().pwh_tuning_set(self.h, _prec(precision), a.ctypes.data))

    def autotune(self, nstates, precision=FP64, target=0):
        """Tunes H|psi> of the first nstates resident states in this precision (pwh_autotune:
        the projections' split-K block target, the GEMMs' LDS stage depths, the runtime FFT
        plans' radix order); returns (target, ms per H|psi>).  target > 0 sets that target
        without timing."""
        t = C.c_int(int(target))
        ms = C.c_double(0.0)
        _check(lib().pwh_autotune(self.h, int(nstates), _prec(precision), C.byref(t), C.byref(ms)))
        return t.value, ms.value

    def resident_chebyfilt(self, nstates, degree, lb, ub, precision=FP64):
        _check(lib().pwh_resident_chebyfilt(self.h, int(nstates), int(degree), float(lb),
                                            float(ub), _prec(precision)))

    def resident_chfsi(self, nstates, degree=10, niter=1, precision=FP64, lanczos_k=20, bounds_mode=0,
                       v0=None, lb=0.0, ub=0.0):
        """chfsi on the first nstates resident states; the Ritz vectors stay resident."""
        if v0 is None:
            v0 = np.ones(self.ngw, np.complex128)
        v0 = np.array(np.asarray(v0, np.complex128).reshape(-1), order="C")
        o = ChfsiOpts(int(degree), int(niter), _prec(precision), int(lanczos_k), int(bounds_mode),
                      float(lb), float(ub), 0, 0, 0, 0, 0, 0)
        eig = np.zeros(nstates)
        res = np.zeros(nstates)
        _check(lib().pwh_resident_chfsi(self.h, int(nstates), _zptr(v0), C.byref(o), eig, res))
        return eig, res, {f: getattr(o, f) for f, _ in ChfsiOpts._fields_}

    def resident_eval_rho_psi(self, nstates, w, omega, rhoe=None):
        """eval_rho_psi on the first nstates resident states."""
        rho = np.zeros(self.nnr) if rhoe is None else np.array(rhoe, dtype=np.float64)
        charge = np.zeros(nstates)
        ek = C.c_double(0.0)
        _check(lib().pwh_resident_eval_rho_psi(self.h, int(nstates), np.ascontiguousarray(w, dtype=np.float64),
                                               float(omega), rho, C.byref(ek), charge))
        return rho, ek.value, charge

    def resident_eval_enl(self, nstates, focc, wkpt):
        """eval_enl on the first nstates resident states: returns (ENL, fnl [nproj][nstates])."""
        enl = C.c_double()
        fnl = np.zeros((max(self.nproj, 1), nstates), np.complex128)
        _check(lib().pwh_resident_eval_enl(self.h, int(nstates), np.ascontiguousarray(focc, dtype=np.float64),
                                           float(wkpt), _zptr(fnl), C.byref(enl)))
        return enl.value, fnl[:self.nproj]


class Batch:
    """Independent systems whose resident blocks are Chebyshev-filtered together
    (pwh_batch_*): mode 0 one after the other, mode 1 one hipGraph with a branch per
    system, mode 2 grouped (each pass of each H|psi> one launch over all systems)."""

    def __init__(self, systems, nstates):
        self.systems = list(systems)
        self.nstates = [int(n) for n in nstates]
        hs = (C.c_void_p * len(self.systems))(*[x.h.value for x in self.systems])
        h = C.c_void_p()
        _check(lib().pwh_batch_create(hs, np.asarray(self.nstates, np.int32), len(self.systems),
                                      C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pwh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, i, X):
        X = _block(X)
        assert X.shape == (self.nstates[i], self.systems[i].ngw)
        _check(lib().pwh_batch_set(self.h, int(i), _zptr(X)))

    def get(self, i):
        X = np.zeros((self.nstates[i], self.systems[i].ngw), np.complex128)
        _check(lib().pwh_batch_get(self.h, int(i), _zptr(X)))
        return X

    def chebyfilt(self, degree, lb, ub, precision=FP64, mode=1):
        lb = np.ascontiguousarray(lb, np.float64)
        ub = np.ascontiguousarray(ub, np.float64)
        _check(lib().pwh_batch_chebyfilt(self.h, int(degree), lb, ub, _prec(precision), int(mode)))


# --------------------------------------------------------------------------
# SCF steps around the eigensolver (EvalEnergy, xc_PZ, mixers, NormalizeRho)
# --------------------------------------------------------------------------

class Grid:
    """The density G-vectors, local form factors and Kerker sphere (pwh_grid_create)."""

    def __init__(self, nr, n1, n2, n3, g2, sfac, vps, rhops, tpiba2, omega, n123=None, xkg=None):
        init()
        self._keep = dict(n1=np.ascontiguousarray(n1, np.int32), n2=np.ascontiguousarray(n2, np.int32),
                          n3=np.ascontiguousarray(n3, np.int32), g2=np.ascontiguousarray(g2, np.float64),
                          sfac=np.ascontiguousarray(np.asarray(sfac, np.complex128).T),    # species fastest
                          vps=np.ascontiguousarray(np.asarray(vps, np.float64).T),
                          rhops=np.ascontiguousarray(np.asarray(rhops, np.float64).T),
                          n123=np.ascontiguousarray(n123 if n123 is not None else [], np.uint32),
                          xkg=np.ascontiguousarray(xkg if xkg is not None else [], np.float64))
        k = self._keep
        nsp = np.asarray(vps).shape[0] if np.ndim(vps) == 2 else 1
        d = GridDesc(nr1=int(nr[0]), nr2=int(nr[1]), nr3=int(nr[2]), ng=len(k["n1"]),
                     n1=k["n1"].ctypes.data, n2=k["n2"].ctypes.data, n3=k["n3"].ctypes.data,
                     g2=k["g2"].ctypes.data, nsp=nsp, sfac=k["sfac"].ctypes.data, vps=k["vps"].ctypes.data,
                     rhops=k["rhops"].ctypes.data, tpiba2=float(tpiba2), omega=float(omega),
                     ngw=len(k["n123"]), n123=k["n123"].ctypes.data if len(k["n123"]) else None,
                     xkg=k["xkg"].ctypes.data if len(k["xkg"]) else None)
        h = C.c_void_p()
        _check(lib().pwh_grid_create(C.byref(d), C.byref(h)))
        self.h = h
        self.nnr = int(nr[0]) * int(nr[1]) * int(nr[2])

    def close(self):
        if self.h:
            lib().pwh_grid_destroy(self.h)
            self.h = None

    def eval_energy(self, rhoe, ekin=0.0, enl=0.0, esr=0.0, eself=0.0):
        """EvalEnergy.cpp:5-163: returns (V_eff, dict of the energy terms)."""
        r = np.array(rhoe, np.float64)
        e = Energy(ekin=ekin, enl=enl, esr=esr, eself=eself)
        _check(lib().pwh_eval_energy(self.h, r, C.byref(e)))
        return r, {f: getattr(e, f) for f, _ in Energy._fields_}

    def mix_kerker(self, beta, q0, rhoe, rho_old):
        d = C.c_double()
        _check(lib().pwh_mix_kerker(self.h, beta, q0, rhoe, rho_old, C.byref(d)))
        return d.value

    def kerker_prec(self, diff, idx1, beta, q0):
        _check(lib().pwh_kerker_prec(self.h, diff.reshape(-1), idx1, beta, q0))

    def mix_pulay(self, it, rhoe, rho_old, f, mixdim, q0, nel):
        d = C.c_double()
        nd = C.c_double()
        _check(lib().pwh_mix_pulay(self.h, it, rhoe, rho_old.reshape(-1), f.reshape(-1), mixdim, q0, nel,
                                   C.byref(d), C.byref(nd)))
        return d.value


MIXERS = {"simple": 0, "kerker": 1, "anderson": 2, "broyden": 4}


class ScfState:
    """The density, V_eff and mixing history of an SCF loop held in HBM
    (pwh_scf_state_*): each step moves only scalars between host and device."""

    def __init__(self, grid, mixer, mixdim=4):
        init()
        h = C.c_void_p()
        _check(lib().pwh_scf_state_create(grid.h, MIXERS[mixer] if isinstance(mixer, str) else int(mixer),
                                          int(mixdim), C.byref(h)))
        self.h = h
        self.nnr = grid.nnr

    def close(self):
        if self.h:
            lib().pwh_scf_state_destroy(self.h)
            self.h = None

    def set_rho(self, rho):
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        if r.size != self.nnr:
            raise ValueError("rho does not match the grid")
        _check(lib().pwh_scf_state_set_rho(self.h, r))

    def get(self):
        """(density, V_eff) as host arrays."""
        rho = np.zeros(self.nnr)
        veff = np.zeros(self.nnr)
        _check(lib().pwh_scf_state_get(self.h, rho.ctypes.data_as(C.c_void_p), veff.ctypes.data_as(C.c_void_p)))
        return rho, veff

    def density(self, sysm, nstates, w, omega):
        """EvalRhoPsi of the first nstates resident states into the held density: returns
        (ekin, charges)."""
        w = np.ascontiguousarray(w, np.float64)
        ek = C.c_double()
        q = np.zeros(nstates)
        _check(lib().pwh_scf_state_density(sysm.h, self.h, int(nstates), w, float(omega), C.byref(ek), q))
        return ek.value, q

    def mix(self, it, beta=0.3, q0=1.0, alpha=0.25, w0=0.01):
        d = C.c_double()
        _check(lib().pwh_scf_state_mix(self.h, int(it), float(beta), float(q0), float(alpha), float(w0),
                                       C.byref(d)))
        return d.value

    def energy(self, ekin=0.0, enl=0.0, esr=0.0, eself=0.0):
        e = Energy(ekin=ekin, enl=enl, esr=esr, eself=eself)
        _check(lib().pwh_scf_state_energy(self.h, C.byref(e)))
        return {f: getattr(e, f) for f, _ in Energy._fields_}

    def set_vloc(self, sysm):
        _check(lib().pwh_scf_state_set_vloc(sysm.h, self.h))


def xc_pz(rho):
    """xc_PZ.cpp:7-63: returns (vxc, Exc, number of points below -0.1)."""
    r = np.ascontiguousarray(rho, np.float64)
    v = np.zeros_like(r)
    exc = C.c_double()
    nneg = C.c_int()
    _check(lib().pwh_xc_pz(r, v, C.byref(exc), r.size, C.byref(nneg)))
    return v, exc.value, nneg.value


def normalize_rho(rho, omega, nel):
    """NormalizeRho.cpp:7-33 in place; returns the electron-count difference."""
    dd = C.c_double()
    _check(lib().pwh_normalize_rho(rho, rho.size, omega, nel, C.byref(dd)))
    return dd.value


def mix_simple(beta, rhoe, rho_old):
    d = C.c_double()
    _check(lib().pwh_mix_simple(beta, rhoe, rho_old, rhoe.size, C.byref(d)))
    return d.value


def mix_anderson(it, beta, rhoe, rho_old, f, g):
    d = C.c_double()
    _check(lib().pwh_mix_anderson(it, beta, rhoe, rho_old.reshape(-1), f.reshape(-1), g.reshape(-1), rhoe.size,
                                  C.byref(d)))
    return d.value


def mix_broyden(it, alpha, w0, rho, rho_old, f, df, u, a, mixdim):
    d = C.c_double()
    _check(lib().pwh_mix_broyden(it, alpha, w0, rho, rho_old.reshape(-1), f.reshape(-1), df.reshape(-1),
                                 u.reshape(-1), a.reshape(-1), mixdim, rho.size, C.byref(d)))
    return d.value
