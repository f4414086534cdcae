"""The block eigensolvers of the SCF driver on the GPU (davidson_cpu.cpp, lobpcg_cpu.cpp)
against the oracle's restatements, FP64, on the Si8 cell (32^3, 1021 plane waves, 16
states, the reference's start vectors and GenPrec preconditioner with tau = -10 as
SCFLoop.cpp:160-169 passes it).

The GPU orthonormalises the start block by CholeskyQR2 where the reference uses
Householder QR; the two differ by a phase per column, which every later step carries
along, so the eigenvalue estimates and residual norms are compared (1e-10 Ha and 1e-8
absolute: rounding differences over 30 iterations stay below 1e-12 and 3e-10 between two
CPU runs that differ only in that QR) and the blocks through the projector onto their
span (1e-8)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG, oracle_globals

pytestmark = pytest.mark.gpu

SYM_DAV = "_Z12davidson_cpuiPdPCdiiS_iS_"
SYM_LOB = "_Z10lobpcg_cpuiPdPCdiiS_iS_"


@pytest.fixture(scope="module")
def setup(si8, gpu_si8, oracle):
    g = oracle_globals(si8)

    def H(X):
        return oracle.apply_ham_block(g, np.ascontiguousarray(X), si8.vloc, si8.ngw, nthreads=8)

    prec = oracle.gen_prec(si8.gv["xkg"][:si8.ngw], si8.tpiba2, si8.a, -10.0)
    X0 = oracle.start_vectors(si8.ngw, 16)
    return H, prec, X0


def _proj(X, k):
    Q, _ = np.linalg.qr(np.asarray(X)[:k].T)
    return Q @ Q.conj().T


@pytest.mark.parametrize("method", ["lobpcg", "davidson"])
def test_solver_matches_oracle(gpu_si8, setup, oracle, method):
    H, prec, X0 = setup
    ref = getattr(oracle, method)(H, X0, prec)
    ev, X, res, st = getattr(gpu_si8, method)(X0, prec)
    assert st["iterations"] == ref.iterations and st["reason"] == ref.reason
    assert np.abs(ev - ref.evals).max() < 1e-10
    assert np.abs(res - ref.res).max() < 1e-8
    # the lowest 15 states (the 16th opens a 3-fold group) span the same space
    assert np.abs(_proj(X, 15) - _proj(ref.X, 15)).max() < 1e-8
    # the reported progress is the oracle's history
    its = [(a, b, v) for e, a, b, v in st["events"] if e == 1]
    assert [(a, b) for a, b, _ in its] == [(a, b) for a, b, _ in ref.history]
    assert np.abs(np.array([v for *_, v in its]) - [v for *_, v in ref.history]).max() < 1e-8
    n_hpsi = 16 * (ref.iterations + 1) if method == "davidson" else 16 * ref.iterations + 16
    assert st["band_applies"] == n_hpsi


@pytest.mark.parametrize("method", ["lobpcg", "davidson"])
def test_nempty_exit(gpu_si8, setup, oracle, method):
    """nconv >= nstates - NEMPTY holds at the first test when NEMPTY = nstates."""
    H, prec, X0 = setup
    ref = getattr(oracle, method)(H, X0, prec, nempty=16)
    ev, X, res, st = getattr(gpu_si8, method)(X0, prec, nempty=16)
    assert st["reason"] == 1 and st["iterations"] == ref.iterations == 1
    assert np.abs(ev - ref.evals).max() < 1e-10


def test_lobpcg_cholesky_failure_is_reported(gpu_si8, setup):
    """A zero preconditioner leaves W = 0, whose overlap has no Cholesky factor
    (lobpcg_cpu.cpp:140-144)."""
    import pwh
    _, prec, X0 = setup
    with pytest.raises(pwh.PwhError) as e:
        gpu_si8.lobpcg(X0, np.zeros_like(prec), max_iter=3)
    assert e.value.code == pwh.PWH_ERR_LAPACK


@pytest.fixture(scope="module")
def shim_diag(si8):
    G = C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    g = si8.globals()
    keep = dict(ngw=np.array([si8.ngw], np.int32), n123=np.ascontiguousarray(si8.gv["n123"]),
                xkg=np.ascontiguousarray(si8.gv["xkg"]), na=g.na.copy(), lmax=g.l_max.copy(),
                lloc=g.l_loc.copy(), wnl=g.wnl.copy(), eigr=np.ascontiguousarray(g.eigr),
                pkg=np.ascontiguousarray(g.pkg_a))
    for name, val in (("NR1", si8.nr[0]), ("NR2", si8.nr[1]), ("NR3", si8.nr[2]),
                      ("NNR", int(np.prod(si8.nr))), ("NGWX", g.ngwx), ("NSP", 1), ("NAX", g.nax),
                      ("NLMAX", 4), ("NKPT", 1), ("NX", 16), ("NEMPTY", 0), ("MAX_DIAG_ITER", 0)):
        C.c_int.in_dll(G, name).value = int(val)
    C.c_bool.in_dll(G, "USER_MAX_DIAG_ITER").value = False
    C.c_bool.in_dll(G, "DIAG_VERBOSE").value = True
    C.c_double.in_dll(G, "TPIBA2").value = float(g.tpiba2)
    for name, arr in (("NGW", "ngw"), ("N123", "n123"), ("XKG", "xkg"), ("NA", "na"),
                      ("L_MAX", "lmax"), ("L_LOC", "lloc"), ("WNL", "wnl"), ("EIGR", "eigr"),
                      ("PKG_A", "pkg")):
        C.c_void_p.in_dll(G, name).value = keep[arr].ctypes.data
    vp, ci = C.c_void_p, C.c_int
    for sym in (SYM_DAV, SYM_LOB):
        getattr(S, sym).argtypes = [ci, vp, vp, ci, ci, vp, ci, vp]
    yield G, S, keep
    del keep


@pytest.mark.parametrize("method,sym", [("lobpcg", SYM_LOB), ("davidson", SYM_DAV)])
def test_shim_solver(si8, shim_diag, setup, oracle, capfd, method, sym):
    """Called as Diagon_*_cpu calls it: the shim makes the reference's start vectors
    itself, prints the reference's lines and leaves eigenvalues and vectors in place."""
    G, S, _ = shim_diag
    H, prec, X0 = setup
    ref = getattr(oracle, method)(H, X0, prec)
    ev = np.zeros(16)
    X = np.zeros((16, si8.ngw), np.complex128)
    rho = np.ascontiguousarray(si8.vloc)
    p = np.ascontiguousarray(prec)
    capfd.readouterr()
    getattr(S, sym)(1, ev.ctypes.data, X.ctypes.data, si8.ngw, 16, rho.ctypes.data, rho.size, p.ctypes.data)
    out = capfd.readouterr().out
    assert np.abs(ev - ref.evals).max() < 1e-10
    assert np.abs(_proj(X, 15) - _proj(ref.X, 15)).max() < 1e-8
    lines = out.splitlines()
    if method == "lobpcg":
        assert lines[0].startswith("Allocated dynamic memory in lobpcg_cpu:")
        assert "Final LOBPCG result" in lines
        assert sum(ln.startswith("iter = ") for ln in lines) == ref.iterations - 1
        tail = lines[lines.index("Final LOBPCG result") + 1:]
        assert len(tail) == 16 and tail[0].split()[0] == "1"
        assert abs(float(tail[0].split()[1]) - ref.evals[0]) < 1e-9
    else:
        assert lines[0].startswith("Allocated dynamic memory in davidson_cpu:")
        assert any(ln.startswith("Final Davidson result nconv=") for ln in lines)
        assert sum(ln.startswith("davidson_cpu: ") for ln in lines) == ref.iterations
        assert lines[-1].startswith("End of Davidson iteration: rnorm =")
