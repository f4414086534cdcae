"""Density and kinetic energy from the wave functions (EvalRhoPsi.cpp:83-157) on the GPU
against the oracle, at FP64: 1e-12 relative on the density, 1e-12 relative on the EKIN
increment, 1e-13 absolute on each state's charge.  At the full 64^3 / 13133 PW / 128 band
size the oracle checks a subset of states, and the whole block is checked through
properties that do not depend on the size: Parseval (sum_r rho = nnr / omega
sum_b w_b charge_b), additivity over disjoint sets of states, and accumulation into rhoe."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG, oracle_globals, relerr

pytestmark = pytest.mark.gpu

SYM_RHO5 = "_Z10EvalRhoPsiiPdPCdii"
SYM_RHO4 = "_Z10EvalRhoPsiiPdPCdi"


def _weights(n, seed):
    rng = np.random.default_rng(seed)
    return 2.0 * rng.uniform(0.2, 1.0, n)


def _check(got, ref):
    rho, ek, q = got
    rr, ekr, qr, bad = ref
    assert bad == -1
    assert relerr(rho, rr) < 1e-12
    assert abs(ek - ekr) < 1e-12 * abs(ekr)
    assert np.abs(q - qr).max() < 1e-13


def test_rho_si8_matches_oracle(si8, gpu_si8, oracle):
    X = si8.psi(16, seed=41)
    w = _weights(16, 1)
    omega = si8.a ** 3
    got = gpu_si8.eval_rho_psi(X, w, omega)
    ref = oracle.eval_rho_psi(oracle_globals(si8), X, si8.ngw, w, omega)
    _check(got, ref)
    # the last plane wave is skipped (EvalRhoPsi.cpp:117): charges are 1 - |c_last|^2
    assert np.abs(got[2] - (1.0 - np.abs(X[:, -1]) ** 2)).max() < 1e-13


def test_rho_padded_ld_and_accumulation(si8, gpu_si8, oracle):
    """ldEigvec > NGW with junk in the padding; rhoe already holds a density."""
    ld = si8.ngw + 37
    X = si8.psi(9, seed=42)
    E = np.full((9, ld), 1e3 + 2e3j)
    E[:, :si8.ngw] = X
    w = _weights(9, 2)
    omega = si8.a ** 3
    rng = np.random.default_rng(3)
    rho0 = rng.random(int(np.prod(si8.nr)))
    got = gpu_si8.eval_rho_psi(E, w, omega, rhoe=rho0)
    ref = oracle.eval_rho_psi(oracle_globals(si8), E, si8.ngw, w, omega, rhoe=rho0)
    _check(got, ref)


def test_rho_single_state_and_zero_weight(si8, gpu_si8):
    """One state of weight zero leaves rhoe unchanged and adds no kinetic energy."""
    X = si8.psi(1, seed=43)
    rho0 = np.linspace(0.0, 1.0, int(np.prod(si8.nr)))
    rho, ek, q = gpu_si8.eval_rho_psi(X, np.zeros(1), 1.0, rhoe=rho0)
    assert np.array_equal(rho, rho0)
    assert ek == 0.0
    assert abs(q[0] - (1.0 - abs(X[0, -1]) ** 2)) < 1e-14


@pytest.mark.parametrize("cell", [2, 9])
def test_rho_small_cells_match_oracle(cell, oracle):
    """36^3 and 40^3 grids (radix 3 and 5 stages, spheres that fill the plane)."""
    import pwh
    from pwh import synth
    s = synth.uneven_batch(count=cell + 1)[cell]
    pwh.init(0)
    g = pwh.System(s.globals())
    nb = s.nbands
    X = s.psi(nb, seed=44 + cell)
    w = _weights(nb, cell)
    got = g.eval_rho_psi(X, w, s.a ** 3)
    ref = oracle.eval_rho_psi(oracle_globals(s), X, s.ngw, w, s.a ** 3)
    _check(got, ref)
    g.close()


def test_rho_si64_subset_and_properties(oracle):
    import pwh
    from pwh import synth
    s = synth.config("si64")
    pwh.init(0)
    g = pwh.System(s.globals())
    X = s.psi(128, seed=45)
    w = _weights(128, 5)
    omega = s.a ** 3
    nnr = int(np.prod(s.nr))
    idx = np.array([0, 1, 2, 63, 64, 100, 126, 127])
    got = g.eval_rho_psi(X[idx], w[idx], omega)
    ref = oracle.eval_rho_psi(oracle_globals(s), X[idx], s.ngw, w[idx], omega)
    _check(got, ref)
    rho, ek, q = g.eval_rho_psi(X, w, omega)
    assert np.abs(q - (1.0 - np.abs(X[:, -1]) ** 2)).max() < 1e-13
    assert abs(rho.sum() - nnr / omega * np.dot(w, q)) < 1e-12 * rho.sum()
    assert rho.min() >= 0.0
    ra, eka, _ = g.eval_rho_psi(X[:64], w[:64], omega)
    rb, ekb, _ = g.eval_rho_psi(X[64:], w[64:], omega, rhoe=ra)
    assert relerr(rb, rho) < 1e-14
    assert abs(eka + ekb - ek) < 1e-13 * ek
    g.close()


def test_rho_rejects_bad_arguments(si8, gpu_si8):
    import pwh
    X = si8.psi(2, seed=46)
    with pytest.raises(pwh.PwhError) as e:
        gpu_si8.eval_rho_psi(X, np.ones(2), 0.0)
    assert e.value.code == pwh.PWH_ERR_ARG
    with pytest.raises(pwh.PwhError) as e:
        gpu_si8.eval_rho_psi(X, np.ones(2), 1.0, ld=si8.ngw - 1)
    assert e.value.code == pwh.PWH_ERR_ARG


@pytest.fixture(scope="module")
def shim_rho(si8):
    """The reference globals for one k-point with NX = 6 states (EvalRhoPsi.cpp reads
    WKPT, FOCC, OMEGA, C0 and EKIN besides the grid and G-vector arrays)."""
    G = C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    g = si8.globals()
    nx = 6
    keep = dict(ngw=np.array([si8.ngw], np.int32), n123=np.ascontiguousarray(si8.gv["n123"]),
                xkg=np.ascontiguousarray(si8.gv["xkg"]), na=g.na.copy(), lmax=g.l_max.copy(),
                lloc=g.l_loc.copy(), wnl=g.wnl.copy(), eigr=np.ascontiguousarray(g.eigr),
                pkg=np.ascontiguousarray(g.pkg_a), wkpt=np.array([0.5]),
                focc=np.array([2.0, 2.0, 2.0, 1.5, 0.5, 0.0]),
                c0=np.full(g.ngwx * nx, 9.0 + 9.0j))
    for name, val in (("NR1", si8.nr[0]), ("NR2", si8.nr[1]), ("NR3", si8.nr[2]),
                      ("NNR", int(np.prod(si8.nr))), ("NGWX", g.ngwx), ("NSP", 1), ("NAX", g.nax),
                      ("NLMAX", 4), ("NKPT", 1), ("NX", nx)):
        C.c_int.in_dll(G, name).value = int(val)
    C.c_double.in_dll(G, "TPIBA2").value = float(g.tpiba2)
    C.c_double.in_dll(G, "OMEGA").value = float(si8.a ** 3)
    for name, arr in (("NGW", "ngw"), ("N123", "n123"), ("XKG", "xkg"), ("NA", "na"),
                      ("L_MAX", "lmax"), ("L_LOC", "lloc"), ("WNL", "wnl"), ("EIGR", "eigr"),
                      ("PKG_A", "pkg"), ("WKPT", "wkpt"), ("FOCC", "focc"), ("C0", "c0")):
        C.c_void_p.in_dll(G, name).value = keep[arr].ctypes.data
    vp, ci = C.c_void_p, C.c_int
    getattr(S, SYM_RHO5).argtypes = [ci, vp, vp, ci, ci]
    getattr(S, SYM_RHO4).argtypes = [ci, vp, vp, ci]
    yield G, S, keep, nx
    del keep


def test_shim_eval_rho_psi(si8, shim_rho, oracle):
    """Both overloads: C0(:, ist, 1) receives the NGWX entries of each state, rhoe gains
    the density, EKIN the kinetic increment (EvalRhoPsi.cpp:5-79, 83-157)."""
    G, S, keep, nx = shim_rho
    ngwx = int(C.c_int.in_dll(G, "NGWX").value)
    omega = si8.a ** 3
    w = keep["wkpt"][0] * keep["focc"]
    nnr = int(np.prod(si8.nr))
    ld = ngwx + 5
    X = si8.psi(nx, seed=47)
    E = np.zeros((nx, ld), np.complex128)
    E[:, :si8.ngw] = X
    rho0 = np.full(nnr, 0.25)
    ref = oracle.eval_rho_psi(oracle_globals(si8), E, si8.ngw, w, omega, rhoe=rho0)
    C.c_double.in_dll(G, "EKIN").value = 1.0
    rho = rho0.copy()
    getattr(S, SYM_RHO5)(1, rho.ctypes.data, E.ctypes.data, ld, si8.ngw)
    assert relerr(rho, ref[0]) < 1e-12
    assert abs(C.c_double.in_dll(G, "EKIN").value - (1.0 + ref[1])) < 1e-12 * (1.0 + ref[1])
    assert np.array_equal(keep["c0"].reshape(nx, ngwx), E[:, :ngwx])
    # four-argument form: ldEigvec = NGWX
    E4 = np.ascontiguousarray(E[:, :ngwx])
    keep["c0"][:] = 9.0
    C.c_double.in_dll(G, "EKIN").value = 0.0
    rho = rho0.copy()
    getattr(S, SYM_RHO4)(1, rho.ctypes.data, E4.ctypes.data, si8.ngw)
    assert relerr(rho, ref[0]) < 1e-12
    assert abs(C.c_double.in_dll(G, "EKIN").value - ref[1]) < 1e-12 * ref[1]
    assert np.array_equal(keep["c0"].reshape(nx, ngwx), E4)
