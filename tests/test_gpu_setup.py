"""The per-G set-up on the GPU (PhaseFactor.cpp, StructureFactor.cpp, NLFormFactor.cpp,
FormFactor.cpp) against the host code and the oracle, FP64, on the silicon
pseudopotential of the reference's example input and the Si8 / Si64 cells.

Tolerances: 1e-14 of the largest value for the phases and structure factor (device sin
and cos against libm) and 1e-12 relative for the radial transforms (a Simpson sum of a
few hundred terms in a different order)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG
from test_oracle_setup import si_pseudo

pytestmark = pytest.mark.gpu


def test_phase_factor_gpu_matches_host(si8):
    import pwh
    pwh.init(0)
    gv = si8.gv
    tau = si8.tau.reshape(8, 3)
    host = pwh.phase_factor(si8.A, si8.nr, [8], 8, tau, gv)
    dev = pwh.phase_factor_gpu(si8.A, si8.nr, [8], 8, tau, gv)
    assert np.abs(dev - host).max() < 1e-14


def test_phase_factor_gpu_odd_grid_and_empty_atom_slots():
    """An odd grid (the reference writes one slot past its table there) and a species
    with fewer atoms than NAX (those EIGR entries are zero)."""
    import pwh
    A = np.diag([9.0, 10.0, 11.0])
    nr = (15, 18, 20)
    gv = pwh.gvectors(A, 6.0, nr, xk=np.array([[0.1, 0.0, 0.05]]))
    tau = np.array([[[0.1, 0.2, 0.3], [1.0, 2.0, 3.0], [0, 0, 0]], [[4.0, 0.5, 1.5], [0, 0, 0], [0, 0, 0]]])
    host = pwh.phase_factor(A, nr, [2, 1], 3, tau, gv)
    dev = pwh.phase_factor_gpu(A, nr, [2, 1], 3, tau, gv)
    assert np.abs(dev - host).max() < 1e-14
    e = dev.reshape(2, 3, gv["ngwx"])
    assert np.all(e[0, 2] == 0) and np.all(e[1, 1:] == 0)


def test_structure_factor_matches_oracle(si8, oracle):
    import pwh
    gv = si8.gv
    tau = si8.tau.reshape(8, 3)
    sf = pwh.structure_factor(si8.A, si8.nr, [8], 8, tau, gv)
    ref = oracle.structure_factor(si8.A, si8.nr, [8], tau, gv["in123"], gv["ng"], gv["ngx"])
    assert np.abs(sf - ref).max() < 1e-13 * np.abs(ref).max()
    assert abs(sf[0, 0] - 8.0) < 1e-13


@pytest.mark.parametrize("cfg,l_loc,xk", [("si8", 3, (0.0, 0.0, 0.0)), ("si8", 1, (0.1, 0.05, 0.0)),
                                          ("si64", 3, (0.0, 0.0, 0.0))])
def test_nl_form_factor_matches_oracle(cfg, l_loc, xk, oracle):
    import pwh
    from pwh import synth
    s = synth.config(cfg)
    ps = si_pseudo(oracle, l_max=3, l_loc=l_loc)
    nlmax = 4 if l_loc == 3 else 9
    ngw = s.ngw if cfg == "si8" else 2000
    xkg = s.gv["xkg"][:ngw]
    ggk = s.gv["ggk"][:3 * ngw]
    omega, tpiba = s.a ** 3, 2 * np.pi / s.a
    wnl, pkg = pwh.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, np.array(xk), nlmax=nlmax)
    wr, pr = oracle.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, np.array(xk), nlmax=nlmax)
    assert np.abs(wnl - wr).max() < 1e-12 * np.abs(wr).max()
    assert np.abs(pkg - pr).max() < 1e-12 * np.abs(pr).max()


def test_form_factor_matches_oracle(si8, oracle):
    import pwh
    from pwh import synth
    ps = si_pseudo(oracle)
    g2 = synth.density_gspace(si8)["g2"]
    omega, tpiba2 = si8.a ** 3, (2 * np.pi / si8.a) ** 2
    vps, rhops, es = pwh.form_factor(ps, [8], omega, tpiba2, g2)
    vr, rr, er = oracle.form_factor(ps, [8], omega, tpiba2, g2)
    assert np.abs(vps - vr).max() < 1e-12 * np.abs(vr).max()
    assert np.abs(rhops - rr).max() < 1e-14 * np.abs(rr).max()
    assert abs(es - er) < 1e-14 * abs(er)


def test_shim_nl_form_factor(si8, oracle):
    """NLFormFactor through its mangled name: WNL, PKG and PKG_A as the reference's globals."""
    import pwh
    G = C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    ps = si_pseudo(oracle)
    d, keep = pwh._pseudo(ps)
    ngw = si8.ngw
    ngwx = si8.globals().ngwx
    xkg = np.zeros(ngwx)
    xkg[:ngw] = si8.gv["xkg"][:ngw]
    ggk = np.zeros(3 * ngwx)
    ggk[:3 * ngw] = si8.gv["ggk"][:3 * ngw]
    xk = np.zeros(3)
    wnl = np.zeros(4)
    pkg = np.zeros(ngwx * 4)
    pkg_a = np.zeros(4 * ngwx)
    ngwa = np.array([ngw], np.int32)
    arrs = dict(MMAX=keep["mmax"], CLOG=keep["clog"], R=keep["r"], PSI=keep["psi"], VION=keep["vion"],
                L_MAX=keep["l_max"], L_LOC=keep["l_loc"], NGW=ngwa, XKG=xkg, GGK=ggk, XK=xk, WNL=wnl, PKG=pkg,
                PKG_A=pkg_a)
    for k, v in arrs.items():
        C.c_void_p.in_dll(G, k).value = v.ctypes.data
    for k, v in (("NSP", 1), ("NKPT", 1), ("NGWX", ngwx), ("NLMAX", 4)):
        C.c_int.in_dll(G, k).value = v
    C.c_double.in_dll(G, "OMEGA").value = si8.a ** 3
    C.c_double.in_dll(G, "TPIBA").value = 2 * np.pi / si8.a
    getattr(S, "_Z12NLFormFactorv")()
    wr, pr = oracle.nl_form_factor(ps, si8.a ** 3, 2 * np.pi / si8.a, ngw, xkg, ggk, xk)
    assert np.abs(wnl - wr[0]).max() < 1e-12 * np.abs(wr).max()
    assert np.abs(pkg_a.reshape(ngwx, 4)[:ngw] - pr[0]).max() < 1e-12 * np.abs(pr).max()
    assert np.array_equal(pkg.reshape(4, ngwx)[:, :ngw], pkg_a.reshape(ngwx, 4)[:ngw].T)
