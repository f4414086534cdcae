"""Pinning of the oracle's form-factor and structure-factor restatements (NLFormFactor.cpp,
FormFactor.cpp, StructureFactor.cpp, simpson.cpp), on the CPU.

* Against the reference's own sources compiled in oracle/_ref, on the silicon
  pseudopotential of the reference's example input (tests/golden/si_fhi.npz, made by
  tests/golden/make_si_fhi.py) and the Si8 cell's G-vectors.
* Against independent quadrature (scipy quad with scipy's spherical Bessel functions) for
  Gaussian radial functions on a logarithmic mesh, so a wrong Bessel combination, radial
  power or prefactor fails.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _ref():
    from oracle import oracle as O
    if not O.Reference.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return O.Reference()


def si_pseudo(oracle, l_max=3, l_loc=3, rgauss=1.0):
    d = np.load(os.path.join(HERE, "golden", "si_fhi.npz"))
    nang = min(int(d["nang"]), 3)
    return oracle.Pseudo(mmax=np.array([int(d["mmax"][0])], np.int32),
                         clog=np.array([np.log(float(d["ratio"][0]))]), r=d["r"][:1].copy(),
                         psi=d["psi"][None, :nang], vion=d["vion"][None, :nang], zv=np.array([float(d["zv"])]),
                         rgauss=np.array([rgauss]), l_max=np.array([l_max], np.int32),
                         l_loc=np.array([l_loc], np.int32))


def test_simpson_is_exact_for_cubics(oracle):
    h = 0.01
    x = np.arange(101) * h
    assert abs(oracle.simpson(x ** 3 - 2 * x, h) - (1.0 ** 4 / 4 - 1.0)) < 1e-14


def test_nl_form_factor_matches_reference(si8, oracle):
    R = _ref()
    ps = si_pseudo(oracle, l_max=3, l_loc=3)
    ngw = si8.ngw
    xkg = si8.gv["xkg"][:ngw]
    ggk = si8.gv["ggk"][:3 * ngw]
    omega, tpiba = si8.a ** 3, 2 * np.pi / si8.a
    wnl, pkg = oracle.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, np.zeros(3))
    wr, pr = R.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, np.zeros(3))
    assert np.abs(wnl - wr).max() < 1e-13 * np.abs(wr).max()
    assert np.abs(pkg - pr).max() < 1e-13 * np.abs(pr).max()
    # s and p channels: four projectors, p along the G direction
    assert np.abs(pkg[0, 0, 1:]).max() == 0.0 and pkg[0, 0, 0] != 0.0


def test_nl_form_factor_d_channel_matches_reference(si8, oracle):
    """L_LOC = 1 makes the p (3) and d (5) channels nonlocal: nine components."""
    R = _ref()
    ps = si_pseudo(oracle, l_max=3, l_loc=1)
    ngw = 300
    xkg = si8.gv["xkg"][:ngw]
    ggk = si8.gv["ggk"][:3 * ngw]
    omega, tpiba = si8.a ** 3, 2 * np.pi / si8.a
    xk = np.array([0.1, 0.05, 0.0])
    wnl, pkg = oracle.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, xk, nlmax=9)
    wr, pr = R.nl_form_factor(ps, omega, tpiba, ngw, xkg, ggk, xk, nlmax=9)
    assert np.abs(wnl - wr).max() < 1e-13 * np.abs(wr).max()
    assert np.abs(pkg - pr).max() < 1e-13 * np.abs(pr).max()


def test_nl_form_factor_against_quadrature(oracle):
    """dv = e^{-r^2/2}, u = r e^{-r^2/2}: PKG_s(q) = int r^2 e^{-r^2} j0(qr) dr, PKG_p(q) along
    G = int r^2 e^{-r^2} j1(qr) dr, WNL = (2l - 1) 4 pi / Omega / int r^2 e^{-3 r^2 / 2} dr."""
    from scipy.integrate import quad
    from scipy.special import spherical_jn
    ratio = 1.0247
    m = 601
    r = 1e-4 * ratio ** np.arange(m)
    gauss = np.exp(-r * r / 2)
    ps = oracle.Pseudo(mmax=np.array([m], np.int32), clog=np.array([np.log(ratio)]), r=r[None, :],
                       psi=np.stack([r * gauss, r * gauss, np.zeros(m)])[None], vion=np.stack(
                           [gauss, gauss, np.zeros(m)])[None], zv=np.array([4.0]), rgauss=np.array([1.0]),
                       l_max=np.array([3], np.int32), l_loc=np.array([3], np.int32))
    q = np.array([0.0, 0.5, 1.3, 2.7])
    tpiba = 1.0
    xkg = q * q
    ggk = np.stack([q, np.zeros(4), np.zeros(4)], axis=1)      # G along x
    omega = 100.0
    wnl, pkg = oracle.nl_form_factor(ps, omega, tpiba, 4, xkg, ggk, np.zeros(3))
    for k, qq in enumerate(q):
        s_ex = quad(lambda x: x * x * np.exp(-x * x) * spherical_jn(0, qq * x), 0, 12, limit=200)[0]
        assert abs(pkg[0, k, 0] - s_ex) < 1e-7
        if qq > 0:
            p_ex = quad(lambda x: x * x * np.exp(-x * x) * spherical_jn(1, qq * x), 0, 12, limit=200)[0]
            assert abs(pkg[0, k, 1] - p_ex) < 1e-7 and abs(pkg[0, k, 2]) < 1e-15
    den = quad(lambda x: x * x * np.exp(-1.5 * x * x), 0, 12)[0]
    assert abs(wnl[0, 0] - 4 * np.pi / omega / den) < 1e-6 * wnl[0, 0]
    assert abs(wnl[0, 1] - 3 * 4 * np.pi / omega / den) < 1e-6 * wnl[0, 1]


def test_form_factor_matches_reference(si8, oracle):
    R = _ref()
    from pwh import synth
    ps = si_pseudo(oracle)
    g2 = synth.density_gspace(si8)["g2"]
    omega, tpiba2 = si8.a ** 3, (2 * np.pi / si8.a) ** 2
    vps, rhops, es = oracle.form_factor(ps, [8], omega, tpiba2, g2)
    vr, rr, er = R.form_factor(ps, [8], omega, tpiba2, g2)
    assert np.abs(vps - vr).max() < 1e-13 * np.abs(vr).max()
    assert np.abs(rhops - rr).max() < 1e-15 * np.abs(rr).max()
    assert abs(es - er) < 1e-14 * abs(er)
    # the pseudo-charge is a normalised Gaussian of charge -ZV: RHOPS(G) -> -ZV/Omega
    assert abs(rhops[0, 1] * omega / -4.0 - np.exp(-0.25 * tpiba2 * g2[1])) < 1e-14


def test_structure_factor_matches_reference(si8, oracle):
    R = _ref()
    g = R.setup(si8.A, si8.ecut, si8.nr, [8], 8, si8.tau.reshape(8, 3))
    ng, ngx = g["ng"], g["ngx"]
    sf = oracle.structure_factor(si8.A, si8.nr, [8], si8.tau.reshape(8, 3), g["in123"], ng, ngx)
    sr = R.structure_factor(1, ng)
    assert np.abs(sf - sr).max() < 1e-13 * np.abs(sr).max()
    assert abs(sf[0, 0] - 8.0) < 1e-14               # G = 0: the atom count
