"""CPU FP64 oracle for the reference's H|psi> / ChFSI path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
imports it.

Two layers:
  * liboracle.so (oracle.c): the C restatement of G-vector generation, phases,
    the 3D FFT conventions, ApplyHam_block, chebyfilt and lanczos.
  * numpy (LAPACK) restatements of the dense steps: ortho_qr (zgeqrf+zungqr,
    ortho_qr.cpp:19-26), eig_zheevd ('V','U', eig_zheevd.cpp:22-25), symmetrize
    (lobpcg_cpu.cpp:15-27), the Rayleigh-Ritz block (ChebySCFLoop.cpp:289-311),
    the residual report (ChebySCFLoop.cpp:415-433) and the bound logic
    (ChebySCFLoop.cpp:7-34,176-196,355-368).

Optionally, oracle/_ref/libref.so (built from the reference's own sources by
`make -C oracle ref`) is wrapped by `Reference` for cross-checks.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so (gcc) and, when the reference tree is present, oracle/_ref."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
        L = C.CDLL(LIB_PATH)
        L.orc_fft_grid.argtypes = [_dp, C.c_double, _ip]
        L.orc_setup_scalars.argtypes = [_dp, C.c_double, _dp]
        L.orc_gvec_sizes.argtypes = [_dp, C.c_double, _ip, _dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_sort1.argtypes = [C.c_int, _dp, _ip]
        L.orc_gvectors.restype = C.c_int
        L.orc_gvectors.argtypes = [_dp, C.c_double, _ip, _dp, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(C.c_int), _ip, _ip, _ip, _dp, _up, _dp]
        L.orc_phase_factor.argtypes = [_dp, _ip, C.c_int, _ip, C.c_int, _dp, C.c_int, _ip, C.c_int,
                                       _ip, C.c_int, _ip, _zp]
        L.orc_fft3d.argtypes = [_zp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_eval_rho_psi.restype = C.c_int
        L.orc_eval_rho_psi.argtypes = [C.c_void_p, C.c_int, _zp, C.c_int, C.c_int, C.c_int, _dp,
                                       C.c_double, _dp, C.POINTER(C.c_double), _dp]
        L.orc_apply_ham_block.restype = C.c_int
        L.orc_apply_ham_block.argtypes = [C.c_void_p, _zp, _zp, C.c_int, C.c_int, _dp, C.c_int, C.c_int]
        L.orc_apply_ham_block_par.restype = C.c_int
        L.orc_apply_ham_block_par.argtypes = [C.c_void_p, _zp, _zp, C.c_int, C.c_int, _dp, C.c_int,
                                              C.c_int, C.c_int]
        L.orc_chebyfilt.restype = C.c_int
        L.orc_chebyfilt.argtypes = [C.c_void_p, C.c_int, _zp, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, _dp, C.c_int]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_void_p, C.c_int, _zp, C.c_int, C.c_int, _dp, C.c_int, _zp, _dp,
                                  C.c_int, C.c_void_p]
        L.orc_induced_norm_mat.restype = C.c_double
        L.orc_induced_norm_mat.argtypes = [_dp, C.c_int, C.c_int]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# Setup restatements
# --------------------------------------------------------------------------

def lattice(A) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(9))


def fft_grid(A, ecut):
    nr = np.zeros(3, np.int32)
    lib().orc_fft_grid(lattice(A), float(ecut), nr)
    return tuple(int(x) for x in nr)


def setup_scalars(A, ecut):
    out = np.zeros(6)
    lib().orc_setup_scalars(lattice(A), float(ecut), out)
    return dict(alat=out[0], omega=out[1], tpiba=out[2], tpiba2=out[3], gcut=out[4], gcutw=out[5])


def sort1(arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    idx = np.zeros(len(arr), np.int32)
    lib().orc_sort1(len(arr), arr, idx)
    return idx


def gvectors(A, ecut, nr, xk=None):
    """Returns the G / G+k maps exactly as the reference's globals hold them."""
    xk = np.zeros((1, 3)) if xk is None else np.asarray(xk, np.float64).reshape(-1, 3)
    nkpt = xk.shape[0]
    xkf = np.ascontiguousarray(xk.reshape(-1))
    nr = np.asarray(nr, np.int32)
    ngx, ngwx = C.c_int(), C.c_int()
    L = lib()
    L.orc_gvec_sizes(lattice(A), float(ecut), nr, xkf, nkpt, C.byref(ngx), C.byref(ngwx))
    ngx, ngwx = ngx.value, ngwx.value
    ng = C.c_int()
    in123 = np.zeros(3 * ngx, np.int32)
    ngw = np.zeros(nkpt, np.int32)
    igk = np.zeros(ngwx * nkpt, np.int32)
    xkg = np.zeros(ngwx * nkpt)
    n123 = np.zeros(ngwx * nkpt, np.uint32)
    ggk = np.zeros(3 * ngwx * nkpt)
    rc = L.orc_gvectors(lattice(A), float(ecut), nr, xkf, nkpt, ngx, ngwx, C.byref(ng), in123, ngw,
                        igk, xkg, n123, ggk)
    if rc != 0:
        raise RuntimeError("orc_gvectors: reference abort condition")
    return dict(ngx=ngx, ngwx=ngwx, ng=ng.value, in123=in123, ngw=ngw, igk=igk, xkg=xkg,
                n123=n123, ggk=ggk, nkpt=nkpt)


def phase_factor(A, nr, na, nax, tau, gv):
    """EIGR [ngwx*nax*nsp*nkpt] (PhaseFactor.cpp:5-80). tau: (nsp, nax, 3) bohr."""
    na = np.asarray(na, np.int32)
    nsp = len(na)
    tau = np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1))
    eigr = np.zeros(gv["ngwx"] * nax * nsp * gv["nkpt"], np.complex128)
    lib().orc_phase_factor(lattice(A), np.asarray(nr, np.int32), nsp, na, nax, tau, gv["nkpt"],
                           gv["ngw"], gv["ngwx"], gv["igk"], gv["ngx"], gv["in123"], eigr)
    return eigr


def fft3d(data, nr, inverse):
    """fft_fftw3 conventions: data flat [nr1*nr2*nr3], x fastest."""
    d = np.ascontiguousarray(np.array(data, dtype=np.complex128).reshape(-1))
    lib().orc_fft3d(d, int(nr[0]), int(nr[1]), int(nr[2]), 1 if inverse else 0)
    return d


# --------------------------------------------------------------------------
# The reference's globals for H|psi>
# --------------------------------------------------------------------------

class _OrcSys(C.Structure):
    _fields_ = [("nr1", C.c_int), ("nr2", C.c_int), ("nr3", C.c_int), ("nnr", C.c_int),
                ("ngwx", C.c_int),
                ("n123", C.c_void_p), ("xkg", C.c_void_p), ("tpiba2", C.c_double),
                ("nsp", C.c_int), ("nax", C.c_int), ("nlmax", C.c_int),
                ("na", C.c_void_p), ("l_max", C.c_void_p), ("l_loc", C.c_void_p),
                ("wnl", C.c_void_p), ("eigr", C.c_void_p), ("pkg_a", C.c_void_p)]


@dataclass
class Globals:
    """The reference globals ApplyHam_block reads (common_pspw.h:43-148)."""
    nr: tuple
    ngwx: int
    n123: np.ndarray     # uint32 [ngwx*nkpt], 1-based
    xkg: np.ndarray      # f64 [ngwx*nkpt]
    tpiba2: float
    na: np.ndarray       # int32 [nsp]
    nax: int
    l_max: np.ndarray    # int32 [nsp]
    l_loc: np.ndarray    # int32 [nsp]
    nlmax: int
    wnl: np.ndarray      # f64 [nsp*nlmax], WNL(is,l) at is + nsp*l
    eigr: np.ndarray     # c128 [ngwx*nax*nsp*nkpt]
    pkg_a: np.ndarray    # f64 [nlmax*ngwx*nsp*nkpt]

    def cstruct(self):
        self._keep = [np.ascontiguousarray(a) for a in (self.n123, self.xkg, self.na, self.l_max,
                                                        self.l_loc, self.wnl, self.eigr, self.pkg_a)]
        n123, xkg, na, lmx, llc, wnl, eigr, pkg = self._keep
        s = _OrcSys(int(self.nr[0]), int(self.nr[1]), int(self.nr[2]),
                    int(self.nr[0] * self.nr[1] * self.nr[2]), int(self.ngwx),
                    n123.ctypes.data, xkg.ctypes.data, float(self.tpiba2),
                    int(len(self.na)), int(self.nax), int(self.nlmax),
                    na.ctypes.data, lmx.ctypes.data, llc.ctypes.data, wnl.ctypes.data,
                    eigr.ctypes.data, pkg.ctypes.data)
        self._s = s
        return C.byref(s)

    @property
    def nnr(self):
        return int(self.nr[0] * self.nr[1] * self.nr[2])


def apply_ham_block(g: Globals, vin, rho, ngw, ik=1, nthreads=1):
    """ApplyHam_block.cpp:5-172.  vin: (nstates, ngw) rows = bands (column-major ld=ngw)."""
    vin = np.ascontiguousarray(vin, dtype=np.complex128)
    nst = vin.shape[0] if vin.ndim == 2 else 1
    vout = np.zeros_like(vin)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    if nthreads > 1:
        rc = lib().orc_apply_ham_block_par(g.cstruct(), vin.reshape(-1), vout.reshape(-1), ngw, ik,
                                           rho, len(rho), nst, nthreads)
    else:
        rc = lib().orc_apply_ham_block(g.cstruct(), vin.reshape(-1), vout.reshape(-1), ngw, ik,
                                       rho, len(rho), nst)
    if rc != 0:
        raise RuntimeError("ApplyHam_block: reference abort condition (nnr != NNR or lm_end > 4)")
    return vout


def eval_rho_psi(g: Globals, eigvec, ngw, w, omega, rhoe=None, ik=1):
    """EvalRhoPsi.cpp:83-157.  eigvec: (nstates, ld) rows = states (column-major, ldEigvec = ld);
    w[ist] = WKPT(ik) * FOCC(ist, ik).  Returns (rhoe + density, EKIN increment, charge per
    state, first state whose charge is off by more than 6e-3 or -1)."""
    ev = np.ascontiguousarray(eigvec, dtype=np.complex128)
    nst, ld = ev.shape
    rho = np.zeros(g.nnr) if rhoe is None else np.array(rhoe, dtype=np.float64)
    charge = np.zeros(nst)
    ek = C.c_double(0.0)
    bad = lib().orc_eval_rho_psi(g.cstruct(), ik, ev.reshape(-1), ld, ngw, nst,
                                 np.ascontiguousarray(w, dtype=np.float64), float(omega), rho,
                                 C.byref(ek), charge)
    return rho, ek.value, charge, bad


def chebyfilt(g: Globals, X, degree, lb, ub, rho, ngw, ik=1):
    """chebyfilt.cpp:6-74 on X (nstates, ngw); returns the filtered copy."""
    X = np.array(X, dtype=np.complex128, order="C")
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    rc = lib().orc_chebyfilt(g.cstruct(), ik, X.reshape(-1), ngw, X.shape[0], degree, lb, ub, rho,
                             len(rho))
    if rc != 0:
        raise RuntimeError("chebyfilt: reference abort condition")
    return X


def lanczos(g: Globals, v0, k, rho, ngw, ik=1, want_v=False):
    """lanczos.cpp:8-101. Returns (T[k,k] as the reference fills it, f, v0_normalised[, V])."""
    v0 = np.array(v0, dtype=np.complex128)
    T = np.zeros(k * k)
    f = np.zeros(ngw, np.complex128)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    V = np.zeros((k, ngw), np.complex128) if want_v else None
    rc = lib().orc_lanczos(g.cstruct(), ik, v0, ngw, k, T, k, f, rho, len(rho),
                           V.ctypes.data if want_v else None)
    if rc != 0:
        raise RuntimeError("lanczos: reference abort condition")
    T = T.reshape(k, k).T.copy()   # column-major -> T[i, j]
    return (T, f, v0, V) if want_v else (T, f, v0)


def induced_norm_mat(T, N):
    """ChebySCFLoop.cpp:7-34 (squared norm, ldA = T.shape[0])."""
    Tf = np.asfortranarray(T, dtype=np.float64).reshape(-1, order="F")
    return lib().orc_induced_norm_mat(np.ascontiguousarray(Tf), T.shape[0], N)


# --------------------------------------------------------------------------
# Dense FP64 steps (numpy -> LAPACK)
# --------------------------------------------------------------------------

def ortho_qr(X):
    """ortho_qr.cpp:3-34: Householder QR (zgeqrf) and explicit Q (zungqr).
    X: (nstates, nbasis) rows = columns of the reference's matrix."""
    Q, _ = np.linalg.qr(np.asarray(X).T, mode="reduced")
    return np.ascontiguousarray(Q.T)


def symmetrize(T):
    """lobpcg_cpu.cpp:15-27 -> (T + T^H)/2 on the stored matrix."""
    return 0.5 * (T + T.conj().T)


def eig_zheevd(T):
    """eig_zheevd.cpp:5-34: jobz='V', uplo='U'."""
    return np.linalg.eigh(T, UPLO="U")


def rayleigh_ritz(X, HX):
    """ChebySCFLoop.cpp:289-311 given orthonormal X and HX (rows = bands):
    T = X^H HX, symmetrize, zheevd, X <- X T.  Returns (eigenvalues, rotated X)."""
    Xm, HXm = X.T, HX.T
    T = symmetrize(Xm.conj().T @ HXm)
    w, Q = eig_zheevd(T)
    return w, np.ascontiguousarray((Xm @ Q).T)


def residual_norms(X, HX):
    """ChebySCFLoop.cpp:415-433: R = HX - X (X^H HX)_sym; per-band 2-norms."""
    Xm, HXm = X.T, HX.T
    G = symmetrize(Xm.conj().T @ HXm)
    R = HXm - Xm @ G
    return np.sqrt(np.real(np.sum(R.conj() * R, axis=0)))


# --------------------------------------------------------------------------
# Block eigensolvers of the SCF driver: Davidson and LOBPCG (FP64, rows = bands)
# --------------------------------------------------------------------------

def random_col(nelem, seed):
    """util_random_cpu.cpp:5-25: the start-vector generator.  The seed is advanced by the
    linear congruential step a <- ((214013 a + 2531011) >> 16) & 32767 and each part is
    (a % 100 + 1) / a with the advanced a, real part first."""
    out = np.empty(nelem, np.complex128)
    a = int(seed)
    for i in range(nelem):
        a = ((a * 214013 + 2531011) >> 16) & 32767
        re = (a % 100 + 1) / a
        a = ((a * 214013 + 2531011) >> 16) & 32767
        im = (a % 100 + 1) / a
        out[i] = complex(re, im)
    return out


def start_vectors(nbasis, nstates):
    """davidson_cpu.cpp:47-51 / lobpcg_cpu.cpp:80-84: column ist from random_col with seed
    100 + ist - 1.  Rows = bands."""
    return np.stack([random_col(nbasis, 100 + i) for i in range(nstates)])


def gen_prec(xkg, tpiba2, alat, tau):
    """GenPrec.cpp:5-27: the Teter-like kinetic preconditioner (tau <= 0) or tau / T(G)."""
    xkg = np.asarray(xkg, np.float64)
    if tau <= 0.0:
        x = xkg * (2.0 * np.pi / alat) ** 2
        y = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
        return y / (y + 16.0 * x ** 4)
    return tau / (0.5 * (xkg + 1.0) * tpiba2)


ZHEGV_SMALL = 12.220446049250313e-16   # eig_zhegv.cpp:8


def eig_zhegv_eval(A, B):
    """eig_zhegv.cpp:82-149 (and eig_zhegv, :5-79, which drops the eigenvalues): canonical
    orthogonalisation.  B = V diag(b) V^H (zheevd, upper triangle); columns with b > SMALL
    scaled by 1/sqrt(b), the others kept; A' = V^H A V; A' = Z diag(e) Z^H (zheevd, upper
    triangle); eigenvectors V Z.  Returns (e, V Z, number of b > SMALL, the small b)."""
    b, V = np.linalg.eigh(B, UPLO="U")
    keep = b > ZHEGV_SMALL
    V = V.copy()
    V[:, keep] /= np.sqrt(b[keep])
    E = A @ V
    A2 = V.conj().T @ E
    e, Z = np.linalg.eigh(A2, UPLO="U")
    return e, V @ Z, int(keep.sum()), b[~keep].copy()


def _cn(X, Y):
    """X^H Y for blocks stored as rows (column-major nbasis x n matrices)."""
    return X.conj() @ Y.T


def _nn(X, M):
    """X M (X rows = columns) -> rows."""
    return M.T @ X


@dataclass
class DiagResult:
    evals: np.ndarray
    X: np.ndarray
    res: np.ndarray
    iterations: int
    nconv: int
    reason: int          # 0: iterations exhausted, 1: nconv >= nstates - nempty, 2: rnorm
    history: list        # (iteration, nconv, rnorm) at the top of each iteration
    evals_history: list  # eigenvalue estimates after each iteration


def davidson(H, X0, prec, nempty=0, max_iter=None, ortho=None):
    """davidson_cpu.cpp:9-247 with H(X) the block H|psi> (rows = bands) and X0 the start
    vectors (before orthonormalisation).  Tolerance 1e-6 on each residual norm, stop when
    nconv >= nstates - nempty or the mean residual < 1e-7."""
    n = X0.shape[0]
    ortho = ortho_qr if ortho is None else ortho
    X = ortho(np.asarray(X0, np.complex128))
    HX = H(X)
    evals = np.real(np.sum(X.conj() * HX, axis=1))
    R = evals[:, None] * X - HX
    res = np.sqrt(np.real(np.sum(R.conj() * R, axis=1)))
    if max_iter is None:
        max_iter = max(30, n // 2)
    mz, tol = 2.220446049250313e-16, 1.0e-6
    rnorm, nconv, reason, hist, it, eh = 1.0, 0, 0, [], 0, []
    for it in range(1, max_iter + 1):
        hist.append((it, nconv, rnorm))
        sc = np.where(res > mz, 1.0 / np.where(res > mz, res, 1.0), 1.0)
        R = (R * sc[:, None]) * prec[None, :]
        HR = H(R)
        h = np.zeros((2 * n, 2 * n), np.complex128)
        o = np.zeros((2 * n, 2 * n), np.complex128)
        h[:n, :n] = _cn(X, HX) if it == 1 else np.diag(evals)
        h[:n, n:] = _cn(X, HR)
        h[n:, :n] = _cn(HR, X)
        h[n:, n:] = _cn(R, HR)
        o[:n, :n] = np.eye(n)
        o[:n, n:] = _cn(X, R)
        o[n:, :n] = _cn(R, X)
        o[n:, n:] = _cn(R, R)
        er, E, _, _ = eig_zhegv_eval(h, o)
        evals = er[:n].copy()
        eh.append(evals)
        X = _nn(X, E[:n, :n]) + _nn(R, E[n:, :n])
        HX = _nn(HX, E[:n, :n]) + _nn(HR, E[n:, :n])
        R = evals[:, None] * X - HX
        res = np.sqrt(np.real(np.sum(R.conj() * R, axis=1)))
        nconv = int(np.sum(res < tol))
        if nconv >= n - nempty:
            reason = 1
            break
        rnorm = float(np.mean(res))
        if rnorm < tol / 10.0:
            reason = 2
            break
    return DiagResult(evals, X, res, it, nconv, reason, hist, eh)


def lobpcg(H, X0, prec, nempty=0, max_iter=None, ortho=None):
    """lobpcg_cpu.cpp:32-389 with H(X) the block H|psi> (rows = bands).  Q = [X | W | P];
    the first iteration uses [X | W] only (:93-181); later ones the full block (:186-367).
    Returns lambda from the last X^H HX, the final X, and the residual norms of that X."""
    n = X0.shape[0]
    ortho = ortho_qr if ortho is None else ortho
    tol, tfudge = 1.0e-6, 1.0e10
    if max_iter is None:
        max_iter = max(30, n // 2)
    X = ortho(np.asarray(X0, np.complex128))
    HX = H(X)
    P = HP = None
    lam = res = None
    nconv, reason, hist, it, eh = 0, 0, [], 0, []
    rnorm = 0.0

    def chol_upper(Cm):
        Cm = symmetrize(Cm)
        L = np.linalg.cholesky(Cm)        # Cm = L L^H, U = L^H (zpotrf 'U')
        return L.conj().T

    def right_solve(Y, U):
        # Y U^{-1} for Y rows = columns: (U^{-T} applied to the row stack)
        return np.linalg.solve(U.T, Y)

    for it in range(1, max(1, max_iter) + 1):
        if it > 1:
            hist.append((it, nconv, rnorm))
        t1 = _cn(X, HX)
        W = HX - _nn(X, t1)
        lam = np.linalg.eigh(t1, UPLO="U")[0]
        eh.append(lam)
        res = np.sqrt(np.real(np.sum(W.conj() * W, axis=1)))
        nconv = int(np.sum(res < tol))
        rnorm = float(np.mean(res))
        if nconv >= n - nempty:
            reason = 1
            break
        if rnorm < tol:
            reason = 2
            break
        W = W * prec[None, :]
        HW = H(W)
        U = chol_upper(_cn(W, W))
        W, HW = right_solve(W, U), right_solve(HW, U)
        if it == 1:
            Q, HQ = np.concatenate([X, W]), np.concatenate([HX, HW])
        else:
            Q, HQ = np.concatenate([X, W, P]), np.concatenate([HX, HW, HP])
        T = symmetrize(_cn(Q, HQ))
        G = symmetrize(_cn(Q, Q))
        _, Uv, _, _ = eig_zhegv_eval(T, G)
        X, HX = _nn(Q, Uv[:, :n]), _nn(HQ, Uv[:, :n])
        if it == 1:
            P, HP = W.copy(), HW.copy()
        else:
            P = _nn(W, Uv[n:2 * n, :n]) + _nn(P, Uv[2 * n:, :n])
            HP = _nn(HW, Uv[n:2 * n, :n]) + _nn(HP, Uv[2 * n:, :n])
            U = chol_upper(_cn(P, P))
            P, HP = right_solve(P, U), right_solve(HP, U)
    return DiagResult(lam, X, res, it, nconv, reason, hist, eh)


# --------------------------------------------------------------------------
# Energies, exchange-correlation and density mixing (FP64; arrays flat, the reference's
# column-major history arrays as (columns, nnr) rows; functions update them in place)
# --------------------------------------------------------------------------

def _seqsum(x):
    """Left-to-right sum, as the reference's loops accumulate."""
    x = np.asarray(x, np.float64)
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def xc_pz(rho):
    """xc_PZ.cpp:7-63: Perdew-Zunger LDA.  Returns (vxc, Exc, indices with rho < -0.1).
    Below -0.1 the reference warns and reuses the previous point's ux, uc, ex, ec (its
    loop-carried locals); before any assignment they are taken as 0."""
    A, B, Cc = -0.1423, 1.0529, 0.3334
    a, b, c, d = 0.0311, -0.0480, 0.0020, -0.0116
    rho = np.asarray(rho, np.float64)
    n = rho.size
    ux = np.zeros(n); uc = np.zeros(n); ex = np.zeros(n); ec = np.zeros(n)
    neg = rho < -0.1
    reg = (~neg) & (rho > 1.0e-13)
    ax = -(3.0 / 4.0) * (3.0 / 2.0 / np.pi) ** (2.0 / 3.0)
    r = rho[reg]
    rs = np.power(3.0 / (4.0 * np.pi * r), 1.0 / 3.0)
    rrs = np.sqrt(rs)
    exr = ax / rs
    uxr = 4.0 / 3.0 * exr
    hi = rs >= 1.0
    bpr = 1.0 + B * rrs + Cc * rs
    ec_hi = A / bpr
    uc_hi = (1.0 + 7.0 / 6.0 * B * rrs + 4.0 / 3.0 * Cc * rs) * ec_hi / bpr
    with np.errstate(divide="ignore", invalid="ignore"):
        rsl = np.log(rs)
    ec_lo = a * rsl + b + c * rs * rsl + d * rs
    uc_lo = (2.0 * d - c) / 3.0 * rs + b - a / 3.0 + a * rsl + 2.0 * c / 3.0 * rs * rsl
    ex[reg] = exr
    ux[reg] = uxr
    ec[reg] = np.where(hi, ec_hi, ec_lo)
    uc[reg] = np.where(hi, uc_hi, uc_lo)
    bad = np.nonzero(neg)[0]
    for i in bad:                       # carry the previous point's values forward
        if i > 0:
            ux[i], uc[i], ex[i], ec[i] = ux[i - 1], uc[i - 1], ex[i - 1], ec[i - 1]
    vxc = ux + uc
    exc = _seqsum(rho * (ex + ec))
    return vxc, exc, bad


@dataclass
class GSpace:
    """The density G-vectors and form factors EvalEnergy reads (GVectorsGen.cpp,
    StructureFactor.cpp, FormFactor.cpp): n1, n2, n3 1-based grid indices [ng], g2 = G
    [ng] in (2pi/a)^2, sfac [nsp][ng] complex, vps and rhops [nsp][ng] real."""
    nr: tuple
    n1: np.ndarray
    n2: np.ndarray
    n3: np.ndarray
    g2: np.ndarray
    sfac: np.ndarray
    vps: np.ndarray
    rhops: np.ndarray
    tpiba2: float
    omega: float

    def gidx(self):
        nr1, nr2, _ = self.nr
        return (self.n1 - 1) + nr1 * ((self.n2 - 1) + nr2 * (self.n3 - 1))


def eval_energy(gs: GSpace, rhoe, ekin, enl, esr, eself):
    """EvalEnergy.cpp:5-163.  Returns (V_eff [nnr], dict of ETOT, EHT, EPSEU, EXC and the
    electron-only / pseudo-charge Hartree sums).  The forward transform of the density is
    kept at grid points outside the G set, so its components there reach V_eff (:113-117)."""
    nr = gs.nr
    rho = np.asarray(rhoe, np.float64)
    cf = fft3d(rho.astype(np.complex128), nr, inverse=False).reshape(-1)
    nsp, ng = gs.sfac.shape
    vp = np.zeros(ng, np.complex128)
    rp = np.zeros(ng, np.complex128)
    for i in range(nsp):                # species summed in order (:51-53, :62-65)
        vp = vp + gs.sfac[i] * gs.vps[i]
        rp = rp + gs.sfac[i] * gs.rhops[i]
    gi = gs.gidx()
    rhet = cf[gi[1:]]
    rhog = rhet + rp[1:]
    fpibg = 4.0 * np.pi / (gs.tpiba2 * gs.g2[1:])
    vcg = fpibg * rhog
    terms_eh = 0.5 * vcg * np.conj(rhog)
    terms_eps = np.conj(rhet) * vp[1:]
    eh = complex(np.cumsum(terms_eh)[-1]) if ng > 1 else 0j
    eps = vp[0] * np.conj(cf[0]) + (complex(np.cumsum(terms_eps)[-1]) if ng > 1 else 0j)
    eh2 = _seqsum(np.real(0.5 * fpibg * rhet * np.conj(rhet)))
    eg2 = _seqsum(np.real(0.5 * fpibg * rp[1:] * np.conj(rp[1:])))
    vtemp = np.concatenate([[vp[0]], vcg + vp[1:]])
    cf[gi] = vtemp
    v = fft3d(cf.reshape(nr[::-1]), nr, inverse=True).reshape(-1)
    vxc, exc, bad = xc_pz(rho)
    nnr = rho.size
    veff = np.real(v) + vxc
    exc = exc * gs.omega / nnr
    eht = np.real(eh) * gs.omega + esr - eself
    epseu = np.real(eps) * gs.omega
    etot = ekin + eht + epseu + enl + exc
    return veff, dict(etot=etot, eht=eht, epseu=epseu, exc=exc, eh_electron=eh2 * gs.omega,
                      eh_pseudo=eg2 * gs.omega, negative=bad)


def normalize_rho(rho, omega, nel):
    """NormalizeRho.cpp:7-33: clip negatives, scale to nel electrons; returns the
    electron-count difference before scaling (the reference warns above 1e-5)."""
    rho = np.asarray(rho, np.float64)
    nnr = rho.size
    truesum = nnr * nel / omega
    r = np.where(rho < 0.0, 0.0, rho)
    s = _seqsum(r)
    if s < 1.0e-15:
        raise RuntimeError("NormalizeRho: sum(rho) < 1e-15")
    rho[:] = r * truesum / s
    return (truesum - s) * omega / nnr


def simple_mix(beta, rhoe, rho_old):
    """SimpleRhoMix.cpp:5-20 (in place); returns d."""
    diff = rhoe - rho_old
    d = np.sqrt(_seqsum(diff * diff) / rhoe.size)
    rhoe[:] = beta * rhoe + (1.0 - beta) * rho_old
    rho_old[:] = rhoe
    return d


def _kerker_filter(vec, nr, beta, q0, n123, xkg):
    cf = fft3d(np.asarray(vec, np.float64).astype(np.complex128), nr, inverse=False).reshape(-1)
    idx = np.asarray(n123, np.int64) - 1
    gg = np.asarray(xkg, np.float64)
    cf[idx] = (beta * gg / (gg + q0)) * cf[idx]
    return np.real(fft3d(cf.reshape(nr[::-1]), nr, inverse=True).reshape(-1))


def kerker_mix(beta, q0, rhoe, rho_old, nr, n123, xkg):
    """KerkerRhoMix.cpp:5-55 with k-point 1's sphere (n123 = N123(1:NGW(1),1), xkg the
    matching XKG): components of rhoe - rho_old on the sphere scaled by beta gg/(gg + q0),
    the others kept (in place); returns d."""
    diff = rhoe - rho_old
    filt = _kerker_filter(diff, nr, beta, q0, n123, xkg)
    d = np.sqrt(_seqsum(diff * diff) / rhoe.size)
    rhoe[:] = rho_old + filt
    rho_old[:] = rhoe
    return d


def kerker_prec(diff_rhoe, idx1, beta, q0, nr, n123, xkg):
    """KerkerRhoMix.cpp:57-82 on column idx1 (1-based) of diff_rhoe (columns, nnr)."""
    diff_rhoe[idx1 - 1] = _kerker_filter(diff_rhoe[idx1 - 1], nr, beta, q0, n123, xkg)


def anderson_mix(it, beta, rhoe, rho_old, f, g):
    """AndersonRhoMix.cpp:9-85: rho_old, f (3, nnr), g (2, nnr), in place; returns d."""
    if it == 1:
        rho_old[0] = rhoe
        f[0] = 0.0
        return 1.0
    j0, j1, j2 = (it - 1) % 3, (it - 2) % 3, (it - 3) % 3
    f[j0] = rhoe - rho_old[j1]
    d = _seqsum(f[j0] * f[j0])
    if it == 2:
        rhoe[:] = beta * rhoe + (1.0 - beta) * rho_old[j1]
        rho_old[j0] = rhoe
        return d
    g[0] = f[j1] - f[j0]
    d11 = float(np.dot(g[0], g[0]))
    r1 = float(np.dot(f[j0], g[0]))
    if it == 3:
        t1 = -r1 / d11
        rhoe[:] = (1.0 - t1) * rho_old[j1] + t1 * rho_old[j2] + beta * (f[j0] + t1 * g[0])
    else:
        g[1] = f[j2] - f[j0]
        d22 = float(np.dot(g[1], g[1]))
        d12 = float(np.dot(g[0], g[1]))
        r2 = float(np.dot(f[j0], g[1]))
        t1 = d11 * d22
        t2 = t1 - d12 * d12
        if abs(t2) > abs(t1) * 1.0e-8:
            t1, t2 = (-r1 * d22 + r2 * d12) / t2, (r1 * d12 - r2 * d11) / t2
        else:
            t1, t2 = 0.0, 0.0
        # the reference adds t2 g(:,1), not t2 g(:,2) (:76-78)
        rhoe[:] = ((1.0 - t1 - t2) * rho_old[j1] + t1 * rho_old[j2] + t2 * rho_old[j0]
                   + beta * (f[j0] + t1 * g[0] + t2 * g[0]))
    rho_old[j0] = rhoe
    return d


def pulay_mix(it, rhoe, rho_old, f, mixdim, nr, n123, xkg, q0, omega, nel):
    """PulayRhoMix.cpp:5-91: rho_old, f (mixdim, nnr) in place; returns d."""
    beta = 0.25
    nnr = rhoe.size
    if mixdim < 2:
        raise RuntimeError("PulayRhoMix: mixdim < 2")
    if it < mixdim:
        rhoe[:] = beta * rhoe + (1.0 - beta) * rho_old[it - 1]
        f[it] = rhoe - rho_old[it - 1]
        kerker_prec(f, it + 1, 1.0, q0, nr, n123, xkg)
        rho_old[it] = rhoe
        return np.sqrt(_seqsum(f[it] * f[it]) / nnr)
    jc = it % mixdim
    jn = (it + 1) % mixdim
    m = min(it + 1, mixdim) + 1
    f[jc] = rhoe - rho_old[jc]
    kerker_prec(f, jc + 1, 1.0, q0, nr, n123, xkg)
    d = np.sqrt(_seqsum(f[jc] * f[jc]) / nnr)
    a = np.zeros((m, m))
    for i in range(m - 1):
        for j in range(i, m - 1):
            a[i, j] = float(np.dot(f[i], f[j]))
        a[i, m - 1] = 1.0
    a = np.triu(a) + np.triu(a, 1).T          # dsysv reads the upper triangle
    rhs = np.zeros(m)
    rhs[m - 1] = 1.0
    alpha = np.linalg.solve(a, rhs)
    new = np.zeros(nnr)
    for j in range(m - 1):
        new = new + alpha[j] * (rho_old[j] + f[j])
    rhoe[:] = new
    normalize_rho(rhoe, omega, nel)
    rho_old[jn] = rhoe
    return d


def broyden_mix(it, alpha, w0, rho, rho_old, f, df, u, a, mixdim):
    """BroydenRhoMix.cpp:10-107: rho_old, f (2, nnr), df, u (mixdim, nnr), a (mixdim x
    mixdim column-major as a[col][row]), all in place; returns d."""
    nnr = rho.size
    if mixdim < 2:
        raise RuntimeError("BroydenRhoMix: mixdim < 2")
    m = min(it + 1, mixdim)
    jc = it % m
    kp = (it - 1) % 2
    kc = it % 2
    f[kc] = rho - rho_old[kp]
    d = np.sqrt(_seqsum(f[kc] * f[kc]) / nnr)
    df[jc] = f[kc] - f[kp]
    t1 = float(np.linalg.norm(df[jc]))
    if t1 > 1.0e-8:
        t1 = 1.0 / t1
    df[jc] = t1 * df[jc]
    u[jc] = alpha * df[jc] + t1 * (rho_old[kp] - rho_old[kc])
    c = np.array([float(np.dot(df[k], f[kc])) for k in range(m)])
    for k in range(m):
        a[jc, k] = float(np.dot(df[jc], df[k]))      # a(k, jc)
        a[k, jc] = a[jc, k]                           # a(jc, k)
    beta = a[:m, :m].T.copy()                        # row-major view of the leading block
    beta = beta + w0 * w0 * np.eye(m)
    binv = np.linalg.inv(beta)
    gamma = np.array([sum(c[k] * binv[k, l] for k in range(m)) for l in range(m)])
    new = rho_old[kp] + alpha * f[kc]
    for l in range(m):
        new = new + (-gamma[l]) * u[l]
    rho[:] = new
    rho_old[kc] = rho
    return d


def eval_enl(g: Globals, c0, ngw, focc, wkpt, ik=1):
    """EvalENL.cpp:5-49 for one k-point: FNL(ist, a, lm) = sum_G C0(G, ist) PKG_A(lm, G)
    conj(EIGR(G, a)) and ENL = WKPT sum FOCC |FNL|^2 WNL(lm).  c0 (nstates, >= ngw) rows =
    states.  Returns (ENL, fnl [nsp][nax][lm][nstates])."""
    c0 = np.asarray(c0, np.complex128)
    nst = c0.shape[0]
    nsp = len(g.na)
    eigr = np.asarray(g.eigr, np.complex128).reshape(nsp, g.nax, g.ngwx)
    pkg = np.asarray(g.pkg_a, np.float64).reshape(nsp, g.ngwx, g.nlmax)
    wnl = np.asarray(g.wnl, np.float64).reshape(g.nlmax, nsp)
    enl = 0.0
    fnl = np.zeros((nsp, g.nax, g.nlmax, nst), np.complex128)
    for i in range(nsp):
        lm_end = int(g.l_max[i]) ** 2 - 2 * int(g.l_loc[i]) + 1
        for ia in range(int(g.na[i])):
            for lm in range(lm_end):
                t = pkg[i, :ngw, lm] * np.conj(eigr[i, ia, :ngw])
                fn = c0[:, :ngw] @ t
                fnl[i, ia, lm] = fn
                enl_m = _seqsum(np.asarray(focc, np.float64)[:nst] * np.abs(fn) ** 2)
                enl = enl + wkpt * enl_m * wnl[lm, i]
    return enl, fnl


# --------------------------------------------------------------------------
# Form factors and structure factor (set-up of the projectors and local potential)
# --------------------------------------------------------------------------

def simpson(f, h):
    """simpson.cpp:5-17 with dsum.cpp (left-to-right sums): h/3 (4 sum f_odd + 2 sum
    f_even-interior + f_0 + f_(N-1)) for odd N."""
    f = np.asarray(f, np.float64)
    n = f.size
    if n % 2 != 1:
        raise RuntimeError("simpson: N must be odd")
    m = (n - 1) // 2
    s = 4.0 * _seqsum(f[1:2 * m:2]) + 2.0 * _seqsum(f[2:2 * m - 1:2]) + f[0] + f[n - 1]
    return s * h / 3.0


@dataclass
class Pseudo:
    """Radial pseudopotential data per species as ReadPseudopotentials.cpp:26-63 holds it:
    mmax [nsp], clog = log(mesh ratio) [nsp], r [nsp][mmax], psi and vion [nsp][nang][mmax],
    zv, rgauss [nsp], l_max, l_loc [nsp]."""
    mmax: np.ndarray
    clog: np.ndarray
    r: np.ndarray
    psi: np.ndarray
    vion: np.ndarray
    zv: np.ndarray
    rgauss: np.ndarray
    l_max: np.ndarray
    l_loc: np.ndarray


def _bessel_fac(l, x):
    if l == 1:
        return np.sin(x) / x
    if l == 2:
        return (np.sin(x) / x - np.cos(x)) / x
    return ((3.0 / (x * x) - 1.0) * np.sin(x) - 3.0 / x * np.cos(x)) / x


def nl_form_factor(ps: Pseudo, omega, tpiba, ngw, xkg, ggk, xk, nlmax=4):
    """NLFormFactor.cpp:5-133 for one k-point: WNL [nsp][nlmax] (WNL(is, m)) and PKG_A
    [nsp][ngw][nlmax] (PKG_A(m, ig, is)); channels l = 1..L_MAX except L_LOC, m = the
    2l - 1 components in the reference's order (s; p_x, p_y, p_z; d_z2, d_zx, d_zy, d_xy,
    d_x2-y2).  xkg [ngw], ggk [ngw][3], xk [3]."""
    nsp = len(ps.mmax)
    fpibo = 4.0 * np.pi / omega
    wnl = np.zeros((nsp, nlmax))
    pkg = np.zeros((nsp, ngw, nlmax))
    xkg = np.asarray(xkg, np.float64)[:ngw]
    ggk = np.asarray(ggk, np.float64).reshape(-1, 3)[:ngw]
    for i in range(nsp):
        m = int(ps.mmax[i])
        h = float(ps.clog[i])
        r = ps.r[i][:m]
        vref = ps.vion[i][int(ps.l_loc[i]) - 1][:m]
        ml = 0
        for l in range(1, int(ps.l_max[i]) + 1):
            if l == int(ps.l_loc[i]):
                continue
            if l >= 4:
                raise RuntimeError(f"not implemented i_l = {l}")
            dv = ps.vion[i][l - 1][:m] - vref
            fun = dv * ps.psi[i][l - 1][:m] * r * r
            q = simpson(dv * ps.psi[i][l - 1][:m] ** 2 * r, h)
            wnl[i, ml:ml + 2 * l - 1] = (2.0 * l - 1.0) * fpibo / q
            for ig in range(ngw):
                t = np.sqrt(xkg[ig])
                arg = t * tpiba
                small = t < 1.0e-4
                if l == 1:
                    q1 = simpson(fun if small else _bessel_fac(1, arg * r) * fun, h)
                    pkg[i, ig, ml] = q1
                    continue
                if small:
                    pkg[i, ig, ml:ml + 2 * l - 1] = 0.0
                    continue
                q1 = simpson(_bessel_fac(l, arg * r) * fun, h)
                c = (ggk[ig] + np.asarray(xk, np.float64)) / t
                if l == 2:
                    pkg[i, ig, ml:ml + 3] = q1 * c
                else:
                    cx_, cy_, cz_ = c
                    s3 = np.sqrt(3.0)
                    pkg[i, ig, ml:ml + 5] = [q1 * 0.5 * (3.0 * cz_ * cz_ - 1.0), q1 * s3 * cz_ * cx_,
                                             q1 * s3 * cz_ * cy_, q1 * s3 * cx_ * cy_,
                                             q1 * s3 / 2.0 * (cx_ * cx_ - cy_ * cy_)]
            ml += 2 * l - 1
    return wnl, pkg


def form_factor(ps: Pseudo, na, omega, tpiba2, g2):
    """FormFactor.cpp:5-72: VPS [nsp][ng] (the local pseudopotential plus the field of
    the Gaussian pseudo-charges), RHOPS [nsp][ng] (G = 0 left 0), ESELF.  g2 = G [ng]
    sorted by shells (the transform is computed once per distinct |G|^2)."""
    from math import erf
    nsp = len(ps.mmax)
    tpiba = np.sqrt(tpiba2)
    fpibom = 4.0 * np.pi / omega
    g2 = np.asarray(g2, np.float64)
    ng = g2.size
    eself = 0.0
    for i in range(nsp):
        eself = eself + ps.zv[i] * ps.zv[i] / ps.rgauss[i] * na[i]
    eself = eself * (1.0 / np.sqrt(2.0 * np.pi))
    vps = np.zeros((nsp, ng))
    rhops = np.zeros((nsp, ng))
    for i in range(nsp):
        facc = 0.25 * tpiba2 * ps.rgauss[i] ** 2
        rhops[i, 1:] = -ps.zv[i] / omega * np.exp(-facc * g2[1:])
        m = int(ps.mmax[i])
        r = ps.r[i][:m]
        ll = int(ps.l_loc[i])
        if ll < 1:
            raise RuntimeError("ll < 1")
        vv = -ps.zv[i] / r * np.array([erf(x) for x in r / ps.rgauss[i]])
        vloc = ps.vion[i][ll - 1][:m] - vv
        h = float(ps.clog[i])
        vps[i, 0] = simpson(vloc * r ** 3, h) * fpibom
        g_old, q1 = -1.0, 0.0
        for ig in range(1, ng):
            if g2[ig] != g_old:
                g_old = g2[ig]
                arg = r * np.sqrt(g2[ig]) * tpiba
                q1 = simpson(vloc * np.sin(arg) / arg * r ** 3, h)
            vps[i, ig] = q1 * fpibom
    return vps, rhops, eself


def structure_factor(A, nr, na, tau, in123, ng, ngx):
    """StructureFactor.cpp:5-29 with PhaseFactor.cpp:32-61's tables: SFAC [nsp][ng] =
    sum over atoms of EI1 EI2 EI3 for the density G-vectors (IN1..IN3 = in123)."""
    A = np.asarray(A, np.float64).reshape(3, 3)
    xmat = np.stack([A[0], A[1], A[2]], axis=1)          # columns A1, A2, A3
    xmati = np.linalg.inv(xmat)
    nsp = len(na)
    tau = np.asarray(tau, np.float64).reshape(nsp, -1, 3)
    in1, in2, in3 = (np.asarray(in123[k * ngx:k * ngx + ng], np.int64) for k in range(3))
    sfac = np.zeros((nsp, ng), np.complex128)
    for i in range(nsp):
        for a in range(int(na[i])):
            tp = xmati @ tau[i, a]
            ar = -2.0 * np.pi * tp
            e1 = np.cos(in1 * ar[0]) + 1j * np.sin(in1 * ar[0])
            e2 = np.cos(in2 * ar[1]) + 1j * np.sin(in2 * ar[1])
            e3 = np.cos(in3 * ar[2]) + 1j * np.sin(in3 * ar[2])
            sfac[i] = sfac[i] + e1 * e2 * e3
    return sfac


class Reference:
    """ctypes wrapper of oracle/_ref/libref.so (the reference's own compiled setup code)."""

    def __init__(self, path=REF_PATH):
        L = C.CDLL(path)
        L.ref_setup.restype = C.c_int
        L.ref_setup.argtypes = [_dp, C.c_double, _ip, C.c_int, _ip, C.c_int, _dp, C.c_int, _dp]
        L.ref_sizes.argtypes = [_ip]
        L.ref_scalars.argtypes = [_dp]
        L.ref_copy.argtypes = [_ip, _ip, _dp, _up, _dp, _zp]
        L.ref_sort1.argtypes = [C.c_int, _dp, _ip]
        L.ref_ortho_qr.argtypes = [_zp, C.c_int, C.c_int]
        L.ref_eig_zheevd.argtypes = [_zp, C.c_int, _dp, C.c_int]
        L.ref_eig_zhegv.argtypes = [_zp, C.c_int, _zp, C.c_int, _zp, C.c_int, C.c_int]
        L.ref_eig_zhegv_eval.argtypes = [_zp, C.c_int, _zp, C.c_int, _dp, _zp, C.c_int, C.c_int]
        L.ref_random_col.argtypes = [_zp, C.c_int, C.c_long]
        L.ref_nl_form_factor.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _ip, _ip, C.c_double, C.c_double,
                                         C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.ref_form_factor.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _ip, _ip, C.c_double, C.c_double,
                                      C.c_int, _dp, _dp, _dp, C.POINTER(C.c_double)]
        L.ref_structure_factor.argtypes = [_zp]
        L.ref_xc_pz.argtypes = [_dp, _dp, C.POINTER(C.c_double), C.c_int]
        L.ref_simple_mix.argtypes = [C.c_double, _dp, _dp, C.c_int, C.POINTER(C.c_double)]
        L.ref_anderson_mix.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_int,
                                       C.POINTER(C.c_double)]
        L.ref_broyden_mix.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
                                      C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.ref_normalize_rho.argtypes = [_dp, C.c_int, C.c_double, C.c_double]
        L.ref_eval_enl.argtypes = [C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _dp, _zp, _zp, _dp, C.c_double, _dp, _zp, C.POINTER(C.c_double)]
        self.L = L

    @staticmethod
    def available(path=REF_PATH):
        return os.path.exists(path)

    def setup(self, A, ecut, nr, na, nax, tau, xk=None):
        xk = np.zeros((1, 3)) if xk is None else np.asarray(xk, np.float64).reshape(-1, 3)
        nkpt = xk.shape[0]
        na = np.asarray(na, np.int32)
        self.L.ref_setup(lattice(A), float(ecut), np.asarray(nr, np.int32), len(na), na, nax,
                         np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1)), nkpt,
                         np.ascontiguousarray(xk.reshape(-1)))
        sz = np.zeros(3 + nkpt, np.int32)
        self.L.ref_sizes(sz)
        ngx, ngwx, ng = int(sz[0]), int(sz[1]), int(sz[2])
        in123 = np.zeros(3 * ngx, np.int32)
        igk = np.zeros(ngwx * nkpt, np.int32)
        xkg = np.zeros(ngwx * nkpt)
        n123 = np.zeros(ngwx * nkpt, np.uint32)
        ggk = np.zeros(3 * ngwx * nkpt)
        eigr = np.zeros(ngwx * nax * len(na) * nkpt, np.complex128)
        self.L.ref_copy(in123, igk, xkg, n123, ggk, eigr)
        sc = np.zeros(15)
        self.L.ref_scalars(sc)
        return dict(ngx=ngx, ngwx=ngwx, ng=ng, ngw=sz[3:].copy(), in123=in123, igk=igk, xkg=xkg,
                    n123=n123, ggk=ggk, eigr=eigr, scalars=sc, nkpt=nkpt)

    def sort1(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        idx = np.zeros(len(arr), np.int32)
        self.L.ref_sort1(len(arr), arr, idx)
        return idx

    def ortho_qr(self, X):
        """X rows = bands -> column-major (nbasis x nstates) buffer is X itself."""
        Xc = np.array(X, dtype=np.complex128, order="C")
        self.L.ref_ortho_qr(Xc.reshape(-1), Xc.shape[1], Xc.shape[0])
        return Xc

    def eig_zheevd(self, T):
        n = T.shape[0]
        A = np.array(np.asarray(T, np.complex128).T, order="C", copy=True)   # column-major buffer
        w = np.zeros(n)
        self.L.ref_eig_zheevd(A.reshape(-1), n, w, n)
        return w, A.T.copy()

    def eig_zhegv_eval(self, A, B):
        """Returns (eigenvalues, eigenvectors) of the reference's eig_zhegv_eval."""
        n = A.shape[0]
        Ac = np.array(np.asarray(A, np.complex128).T, order="C", copy=True)
        Bc = np.array(np.asarray(B, np.complex128).T, order="C", copy=True)
        E = np.zeros((n, n), np.complex128)
        w = np.zeros(n)
        self.L.ref_eig_zhegv_eval(Ac.reshape(-1), n, Bc.reshape(-1), n, w, E.reshape(-1), n, n)
        return w, E.T.copy()

    def eig_zhegv(self, A, B):
        n = A.shape[0]
        Ac = np.array(np.asarray(A, np.complex128).T, order="C", copy=True)
        Bc = np.array(np.asarray(B, np.complex128).T, order="C", copy=True)
        E = np.zeros((n, n), np.complex128)
        self.L.ref_eig_zhegv(Ac.reshape(-1), n, Bc.reshape(-1), n, E.reshape(-1), n, n)
        return E.T.copy()

    def random_col(self, nelem, seed):
        out = np.zeros(nelem, np.complex128)
        self.L.ref_random_col(out, nelem, seed)
        return out

    def xc_pz(self, rho):
        rho = np.array(rho, np.float64)
        vxc = np.zeros_like(rho)
        exc = C.c_double()
        self.L.ref_xc_pz(rho, vxc, C.byref(exc), rho.size)
        return vxc, exc.value

    def simple_mix(self, beta, rhoe, rho_old):
        d = C.c_double()
        self.L.ref_simple_mix(beta, rhoe, rho_old, rhoe.size, C.byref(d))
        return d.value

    def anderson_mix(self, it, beta, rhoe, rho_old, f, g):
        d = C.c_double()
        self.L.ref_anderson_mix(it, beta, rhoe, rho_old.reshape(-1), f.reshape(-1), g.reshape(-1),
                                rhoe.size, C.byref(d))
        return d.value

    def broyden_mix(self, it, alpha, w0, rho, rho_old, f, df, u, a, mixdim):
        d = C.c_double()
        self.L.ref_broyden_mix(it, alpha, w0, rho, rho_old.reshape(-1), f.reshape(-1), df.reshape(-1),
                               u.reshape(-1), a.reshape(-1), mixdim, rho.size, C.byref(d))
        return d.value

    def normalize_rho(self, rho, omega, nel):
        self.L.ref_normalize_rho(rho, rho.size, omega, nel)

    def eval_enl(self, g, c0, ngw, focc, wkpt):
        c0 = np.ascontiguousarray(c0, np.complex128)
        nst = c0.shape[0]
        nsp = len(g.na)
        fnl = np.zeros(nst * g.nax * nsp * g.nlmax, np.complex128)
        enl = C.c_double()
        self.L.ref_eval_enl(nsp, np.asarray(g.na, np.int32), g.nax, np.asarray(g.l_max, np.int32),
                            np.asarray(g.l_loc, np.int32), g.nlmax, ngw, c0.shape[1], nst,
                            np.ascontiguousarray(g.pkg_a, np.float64), np.ascontiguousarray(g.eigr),
                            c0.reshape(-1), np.ascontiguousarray(focc, np.float64), float(wkpt),
                            np.ascontiguousarray(g.wnl, np.float64), fnl, C.byref(enl))
        # FNL(ist, ik, is, ia, lm) with ik = 1: ist fastest, then is, ia, lm
        return enl.value, fnl.reshape(g.nlmax, g.nax, nsp, nst).transpose(2, 1, 0, 3)

    @staticmethod
    def _radial(ps, maxr=570, maxam=3):
        nsp = len(ps.mmax)
        r = np.zeros(maxr * nsp)
        psi = np.zeros(maxr * nsp * maxam)
        vion = np.zeros(maxr * nsp * maxam)
        for i in range(nsp):
            m = int(ps.mmax[i])
            r[i * maxr:i * maxr + m] = ps.r[i][:m]
            for l in range(min(maxam, ps.psi.shape[1])):
                o = (l * nsp + i) * maxr
                psi[o:o + m] = ps.psi[i][l][:m]
                vion[o:o + m] = ps.vion[i][l][:m]
        return r, psi, vion

    def nl_form_factor(self, ps, omega, tpiba, ngw, xkg, ggk, xk, nlmax=4):
        """Returns (WNL [nsp][nlmax], PKG_A [nsp][ngw][nlmax]) of the reference's NLFormFactor."""
        nsp = len(ps.mmax)
        r, psi, vion = self._radial(ps)
        ngwx = ngw
        wnl = np.zeros(nsp * nlmax)
        pkg = np.zeros(ngwx * nsp * nlmax)
        pkg_a = np.zeros(nlmax * ngwx * nsp)
        self.L.ref_nl_form_factor(nsp, np.asarray(ps.mmax, np.int32), np.asarray(ps.clog, np.float64), r, psi,
                                  vion, np.asarray(ps.l_max, np.int32), np.asarray(ps.l_loc, np.int32), omega,
                                  tpiba, ngw, ngwx, np.ascontiguousarray(xkg[:ngw], np.float64),
                                  np.ascontiguousarray(np.asarray(ggk).reshape(-1)[:3 * ngw], np.float64),
                                  np.ascontiguousarray(xk, np.float64), nlmax, wnl, pkg, pkg_a)
        return wnl.reshape(nlmax, nsp).T.copy(), pkg_a.reshape(nsp, ngwx, nlmax)

    def form_factor(self, ps, na, omega, tpiba2, g2):
        nsp = len(ps.mmax)
        r, psi, vion = self._radial(ps)
        ng = len(g2)
        vps = np.zeros(nsp * ng)
        rhops = np.zeros(nsp * ng)
        es = C.c_double()
        self.L.ref_form_factor(nsp, np.asarray(ps.mmax, np.int32), np.asarray(ps.clog, np.float64), r, vion,
                               np.asarray(ps.zv, np.float64), np.asarray(ps.rgauss, np.float64),
                               np.asarray(na, np.int32), np.asarray(ps.l_loc, np.int32), omega, tpiba2, ng,
                               np.ascontiguousarray(g2, np.float64), vps, rhops, C.byref(es))
        return vps.reshape(ng, nsp).T.copy(), rhops.reshape(ng, nsp).T.copy(), es.value

    def structure_factor(self, nsp, ng):
        """SFAC [nsp][ng] of the reference's StructureFactor on the set-up of setup()."""
        sf = np.zeros(nsp * ng, np.complex128)
        self.L.ref_structure_factor(sf)
        return sf.reshape(ng, nsp).T.copy()
