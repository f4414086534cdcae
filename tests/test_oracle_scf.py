"""Pinning of the oracle's energy, exchange-correlation and density-mixing restatements
(EvalEnergy.cpp, xc_PZ.cpp, *RhoMix.cpp, NormalizeRho.cpp, EvalENL.cpp), on the CPU.

* xc_PZ, SimpleRhoMix, AndersonRhoMix, BroydenRhoMix, NormalizeRho and EvalENL: against
  the reference's own sources compiled in oracle/_ref, over several mixing iterations.
* xc_PZ also against the thermodynamic identity v_xc = d(rho eps_xc)/d rho (finite
  differences) and the closed-form exchange energy -(3/4)(3/pi)^(1/3) rho^(4/3).
* KerkerRhoMix / KerkerPrec (which need the FFT the image cannot build for the
  reference): single plane waves inside and outside the sphere and the G = 0 component.
* PulayRhoMix: its coefficients solve the constrained least-squares problem (an
  independent KKT solve) and the result is normalised.
* EvalEnergy: a cosine density and a Gaussian local form factor in closed form (Hartree
  energy, pseudopotential energy, Hartree potential).
"""
import numpy as np
import pytest

from conftest import oracle_globals


def _ref():
    from oracle import oracle as O
    if not O.Reference.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return O.Reference()


def test_xc_pz_matches_reference(oracle):
    R = _ref()
    rng = np.random.default_rng(1)
    rho = np.concatenate([rng.uniform(0, 2.0, 500), 10.0 ** rng.uniform(-16, 1, 500),
                          [0.0, 1e-13, 2e-13, 3.0 / (4 * np.pi), -0.05]])
    rho[100] = -0.3               # below -0.1: the previous point's values are reused
    rho[101] = -0.2
    vxc, exc, bad = oracle.xc_pz(rho)
    vr, er = R.xc_pz(rho)
    assert list(bad) == [100, 101]
    assert np.abs(vxc - vr).max() <= 4e-16 * np.abs(vr).max()
    assert abs(exc - er) <= 1e-14 * abs(er)


def test_xc_pz_closed_forms(oracle):
    """Exchange: e_x = -(3/4)(3/pi)^(1/3) rho^(1/3) per electron; v_xc = d(rho eps)/d rho on
    both sides of rs = 1 (rho = 3/(4 pi) is rs = 1)."""
    rho = np.array([0.02, 0.1, 0.2, 0.5, 3.0])
    for r in rho:
        h = 1e-6 * r
        ep = oracle.xc_pz(np.array([r + h]))[1]
        em = oracle.xc_pz(np.array([r - h]))[1]
        v = oracle.xc_pz(np.array([r]))[0][0]
        assert abs((ep - em) / (2 * h) - v) < 1e-7 * abs(v)
    # at rho = 1e3 (rs < 1) the total minus the rs < 1 correlation formula is the exchange
    # energy in its closed form
    r = 1e3
    vxc, exc, _ = oracle.xc_pz(np.array([r]))
    rs = (3 / (4 * np.pi * r)) ** (1 / 3)
    ec = 0.0311 * np.log(rs) - 0.0480 + 0.0020 * rs * np.log(rs) - 0.0116 * rs
    ex = -(3 / 4) * (3 / np.pi) ** (1 / 3) * r ** (1 / 3)
    assert abs(exc / r - ec - ex) < 1e-12 * abs(ex)


def test_mixers_match_reference_over_iterations(oracle):
    R = _ref()
    rng = np.random.default_rng(2)
    n = 3000
    seq = [rng.random(n) for _ in range(7)]
    # simple
    a_old, b_old = seq[0].copy(), seq[0].copy()
    for x in seq[1:]:
        a, b = x.copy(), x.copy()
        da = oracle.simple_mix(0.3, a, a_old)
        db = R.simple_mix(0.3, b, b_old)
        assert np.array_equal(a, b) and np.array_equal(a_old, b_old) and abs(da - db) < 1e-15
    # Anderson, iterations 1..6 (covers the iter = 1, 2, 3 and >= 4 branches)
    ro, f, g = np.zeros((3, n)), np.zeros((3, n)), np.zeros((2, n))
    ro2, f2, g2 = ro.copy(), f.copy(), g.copy()
    for it, x in enumerate(seq[:6], start=1):
        a, b = x.copy(), x.copy()
        da = oracle.anderson_mix(it, 0.4, a, ro, f, g)
        db = R.anderson_mix(it, 0.4, b, ro2, f2, g2)
        assert np.abs(a - b).max() < 1e-13 and abs(da - db) <= 1e-12 * abs(db)
        assert np.abs(ro - ro2).max() < 1e-13
    # Broyden, mixdim 4, iterations 1..7
    md = 4
    st = [np.zeros((2, n)), np.zeros((2, n)), np.zeros((md, n)), np.zeros((md, n)), np.zeros((md, md))]
    st2 = [x.copy() for x in st]
    st[0][:] = seq[0]
    st2[0][:] = seq[0]
    for it, x in enumerate(seq, start=1):
        a, b = x.copy(), x.copy()
        da = oracle.broyden_mix(it, 0.25, 0.01, a, *st, md)
        db = R.broyden_mix(it, 0.25, 0.01, b, *st2, md)
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max(), it
        assert abs(da - db) <= 1e-12 * abs(db)


def test_normalize_rho_matches_reference(oracle):
    R = _ref()
    rng = np.random.default_rng(3)
    rho = rng.random(4000) - 0.05
    a, b = rho.copy(), rho.copy()
    oracle.normalize_rho(a, 270.0, 32.0)
    R.normalize_rho(b, 270.0, 32.0)
    assert np.abs(a - b).max() < 1e-15 * np.abs(b).max()
    assert a.min() >= 0.0 and abs(a.sum() * 270.0 / a.size - 32.0) < 1e-11


def test_eval_enl_matches_reference(si8, oracle):
    R = _ref()
    g = oracle_globals(si8)
    rng = np.random.default_rng(4)
    nst = 6
    c0 = np.zeros((nst, g.ngwx), np.complex128)
    c0[:, :si8.ngw] = si8.psi(nst, seed=5)
    focc = rng.uniform(0.0, 2.0, nst)
    enl, fnl = oracle.eval_enl(g, c0, si8.ngw, focc, 0.5)
    er, fr = R.eval_enl(g, c0, si8.ngw, focc, 0.5)
    assert abs(enl - er) < 1e-12 * abs(er)
    assert np.abs(fnl - fr).max() < 1e-13 * np.abs(fr).max()


def _sphere(nr, a, ecut):
    from pwh import gvectors
    A = np.diag([a, a, a])
    gv = gvectors(A, ecut, nr)
    ngw = int(gv["ngw"][0])
    return gv["n123"][:ngw].astype(np.int64), gv["xkg"][:ngw]


def test_kerker_closed_forms(oracle):
    nr = (16, 16, 16)
    n123, xkg = _sphere(nr, 10.0, 4.0)
    x = np.arange(16)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")           # X = i1 (fastest in memory)
    flat = lambda v: np.ascontiguousarray(v.transpose(2, 1, 0)).reshape(-1)
    wave_in = flat(np.cos(2 * np.pi * 2 * X / 16))           # |G|^2 = 4 (2pi/a)^2: on the sphere
    wave_out = flat(np.cos(2 * np.pi * 7 * X / 16))          # |G|^2 = 49: outside
    assert 4.0 in set(np.round(xkg, 12)) and 49.0 not in set(np.round(xkg, 12))
    beta, q0 = 0.7, 1.5
    old = np.full(16 ** 3, 0.3)
    new = old + 0.2 + wave_in + 0.5 * wave_out
    r, o = new.copy(), old.copy()
    d = oracle.kerker_mix(beta, q0, r, o, nr, n123, xkg)
    expect = old + beta * 4.0 / (4.0 + q0) * wave_in + 0.5 * wave_out   # G = 0 removed
    assert np.abs(r - expect).max() < 1e-13
    assert np.array_equal(o, r)
    assert abs(d - np.sqrt(np.mean((new - old) ** 2))) < 1e-14
    diff = np.stack([np.zeros(16 ** 3), wave_in + 1.0])
    oracle.kerker_prec(diff, 2, 1.0, q0, nr, n123, xkg)
    assert np.abs(diff[1] - 4.0 / (4.0 + q0) * wave_in).max() < 1e-13
    assert np.all(diff[0] == 0.0)


def test_pulay_coefficients_solve_the_constrained_least_squares(oracle):
    """In the history branch, rhoe = N(sum_j alpha_j (rho_old_j + f_j)) with alpha minimising
    ||sum alpha_j f_j|| subject to sum alpha_j = 1 (N: NormalizeRho)."""
    nr = (8, 8, 8)
    n123, xkg = _sphere(nr, 6.0, 3.0)
    nnr = 512
    rng = np.random.default_rng(6)
    md = 3
    ro, f = np.zeros((md, nnr)), np.zeros((md, nnr))
    ro[0] = rng.random(nnr) + 1.0
    omega, nel = 216.0, 8.0
    outs = []
    for it in range(1, 6):
        x = rng.random(nnr) + 1.0
        oracle.pulay_mix(it, x, ro, f, md, nr, n123, xkg, 0.5, omega, nel)
        outs.append(x)
    # replay the last call's algebra independently from the state before it
    ro2, f2 = np.zeros((md, nnr)), np.zeros((md, nnr))
    ro2[0] = ro[0]
    rng = np.random.default_rng(6)
    ro2[0] = rng.random(nnr) + 1.0
    for it in range(1, 5):
        x = rng.random(nnr) + 1.0
        oracle.pulay_mix(it, x, ro2, f2, md, nr, n123, xkg, 0.5, omega, nel)
    x = rng.random(nnr) + 1.0
    it = 5
    jc = it % md
    fj = f2.copy()
    fj[jc] = x - ro2[jc]
    oracle.kerker_prec(fj, jc + 1, 1.0, 0.5, nr, n123, xkg)
    Fm = fj[:md]
    K = np.zeros((md + 1, md + 1))
    K[:md, :md] = Fm @ Fm.T
    K[:md, md] = K[md, :md] = 1.0
    alpha = np.linalg.lstsq(K, np.r_[np.zeros(md), 1.0], rcond=None)[0][:md]
    assert abs(alpha.sum() - 1.0) < 1e-12
    raw = sum(alpha[j] * (ro2[j] + fj[j]) for j in range(md))
    oracle.normalize_rho(raw, omega, nel)
    assert np.abs(outs[-1] - raw).max() < 1e-9 * np.abs(raw).max()


def test_eval_energy_closed_form(oracle):
    """Cubic cell a, one species at the origin (S(G) = 1), V_ps(G) = V0 exp(-|G|^2),
    no pseudo-charge; density rho0 + A cos(G0 . r) with G0 = (2pi/a)(2,0,0).  Then
    EHT = Omega pi A^2 / (TPIBA2 |G0|^2), EPSEU = Omega V0 (rho0 + A exp(-|G0|^2)), and
    V_eff - V_xc = (4 pi A / (TPIBA2 |G0|^2)) cos(G0 . r) + sum_G V_ps(G) e^{iG.r}."""
    from pwh import gvectors
    a, nr = 8.0, (16, 16, 16)
    A = np.diag([a, a, a])
    gv = gvectors(A, 3.0, nr)                         # density sphere: 4 x 3 Ry
    ng = gv["ng"]
    ngx = gv["ngx"]
    in1, in2, in3 = (gv["in123"][k * ngx:k * ngx + ng] for k in range(3))
    n1 = np.where(in1 < 0, in1 + 1 + nr[0], in1 + 1)
    n2 = np.where(in2 < 0, in2 + 1 + nr[1], in2 + 1)
    n3 = np.where(in3 < 0, in3 + 1 + nr[2], in3 + 1)
    g2 = (in1 ** 2 + in2 ** 2 + in3 ** 2).astype(np.float64)
    tpiba2 = (2 * np.pi / a) ** 2
    V0 = -0.7
    gs = oracle.GSpace(nr=nr, n1=n1, n2=n2, n3=n3, g2=g2, sfac=np.ones((1, ng), np.complex128),
                       vps=(V0 * np.exp(-g2))[None, :], rhops=np.zeros((1, ng)), tpiba2=tpiba2,
                       omega=a ** 3)
    x = np.arange(16)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rho0, Am = 0.05, 0.02
    rho = np.ascontiguousarray((rho0 + Am * np.cos(2 * np.pi * 2 * X / 16)).transpose(2, 1, 0)).reshape(-1)
    veff, e = oracle.eval_energy(gs, rho, ekin=0.0, enl=0.0, esr=0.0, eself=0.0)
    om = a ** 3
    assert abs(e["eht"] - om * np.pi * Am ** 2 / (tpiba2 * 4.0)) < 1e-12 * abs(e["eht"])
    assert abs(e["epseu"] - om * V0 * (rho0 + Am * np.exp(-4.0))) < 1e-12 * abs(e["epseu"])
    vxc, exc, _ = oracle.xc_pz(rho)
    assert abs(e["exc"] - exc * om / rho.size) < 1e-15 * abs(e["exc"])
    assert abs(e["etot"] - (e["eht"] + e["epseu"] + e["exc"])) < 1e-12
    # potential in real space
    ph = 2 * np.pi * (np.outer(in1, X.transpose(2, 1, 0).reshape(-1)) +
                      np.outer(in2, Y.transpose(2, 1, 0).reshape(-1)) +
                      np.outer(in3, Z.transpose(2, 1, 0).reshape(-1))) / 16
    vps_r = np.real(np.exp(1j * ph).T @ (V0 * np.exp(-g2)))
    vh_r = 4 * np.pi * Am / (tpiba2 * 4.0) * np.cos(2 * np.pi * 2 * X / 16).transpose(2, 1, 0).reshape(-1)
    assert np.abs(veff - vxc - vps_r - vh_r).max() < 1e-12
