"""The SCF steps around the eigensolver on the GPU (EvalEnergy.cpp, xc_PZ.cpp, the density
mixers, NormalizeRho.cpp, EvalENL.cpp) against the oracle, FP64, on the Si8 cell (32^3
grid, 3.3k density G-vectors) with the density of 16 states.

Tolerances: 1e-12 relative on energies, 1e-12 of the largest value on potentials and
mixed densities (sums over the grid run in a different order on the GPU); the Pulay and
Broyden coefficients come from small LU solves (rocSOLVER vs numpy), 1e-10."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG, density_gspace, oracle_globals, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scf(si8, oracle):
    import pwh
    pwh.init(0)
    gs = density_gspace(si8)
    X = si8.psi(16, seed=61)
    rho, _, _, _ = oracle.eval_rho_psi(oracle_globals(si8), X, si8.ngw, np.full(16, 2.0), gs.omega)
    ngw = si8.ngw
    n123 = si8.gv["n123"][:ngw].astype(np.int64)
    xkg = si8.gv["xkg"][:ngw]
    grid = pwh.Grid(si8.nr, gs.n1, gs.n2, gs.n3, gs.g2, gs.sfac, gs.vps, gs.rhops, gs.tpiba2, gs.omega,
                    n123=n123, xkg=xkg)
    yield gs, rho, grid, n123, xkg
    grid.close()


def test_xc_pz_matches_oracle(oracle):
    import pwh
    pwh.init(0)
    rng = np.random.default_rng(7)
    rho = np.concatenate([rng.uniform(0, 2.0, 4000), 10.0 ** rng.uniform(-16, 1, 4000), [0.0, 1e-13]])
    rho[10], rho[11], rho[5000] = -0.3, -0.2, -0.5     # reuse the previous point's terms
    v, exc, nneg = pwh.xc_pz(rho)
    vr, er, bad = oracle.xc_pz(rho)
    assert nneg == 3 and list(bad) == [10, 11, 5000]
    assert np.abs(v - vr).max() < 1e-14 * np.abs(vr).max()
    assert abs(exc - er) < 1e-13 * abs(er)


def test_eval_energy_matches_oracle(scf, oracle):
    gs, rho, grid, _, _ = scf
    ins = dict(ekin=12.5, enl=3.25, esr=-1.5, eself=7.0)
    veff, e = grid.eval_energy(rho, **ins)
    vr, er = oracle.eval_energy(gs, rho, **ins)
    assert np.abs(veff - vr).max() < 1e-12 * np.abs(vr).max()
    for k in ("etot", "eht", "epseu", "exc"):
        assert abs(e[k] - er[k]) < 1e-12 * max(abs(er[k]), 1.0), k
    assert abs(e["eh_electron"] - er["eh_electron"]) < 1e-12 * abs(er["eh_electron"])
    assert abs(e["eh_pseudo"] - er["eh_pseudo"]) < 1e-12 * abs(er["eh_pseudo"])
    assert e["n_negative"] == 0


def test_normalize_rho_matches_oracle(scf, oracle):
    import pwh
    gs, rho, _, _, _ = scf
    a = rho - 0.01
    b = a.copy()
    da = pwh.normalize_rho(a, gs.omega, 30.0)
    db = oracle.normalize_rho(b, gs.omega, 30.0)
    assert np.abs(a - b).max() < 1e-14 * np.abs(b).max()
    assert abs(da - db) < 1e-10
    with pytest.raises(pwh.PwhError):
        pwh.normalize_rho(np.full(100, -1.0), 1.0, 1.0)


def _seq(rho, k, seed):
    rng = np.random.default_rng(seed)
    return [rho * (1.0 + 0.1 * rng.standard_normal(rho.size)) for _ in range(k)]


def test_simple_and_kerker_mixing(scf, oracle):
    import pwh
    gs, rho, grid, n123, xkg = scf
    seq = _seq(rho, 4, 8)
    o1, o2 = rho.copy(), rho.copy()
    for x in seq:
        a, b = x.copy(), x.copy()
        da = pwh.mix_simple(0.3, a, o1)
        db = oracle.simple_mix(0.3, b, o2)
        assert np.abs(a - b).max() < 1e-15 * np.abs(b).max() and abs(da - db) < 1e-14 * db
    o1, o2 = rho.copy(), rho.copy()
    for x in seq:
        a, b = x.copy(), x.copy()
        da = grid.mix_kerker(0.6, 1.2, a, o1)
        db = oracle.kerker_mix(0.6, 1.2, b, o2, gs.nr, n123, xkg)
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max() and abs(da - db) < 1e-12 * db
        assert np.array_equal(a, o1)
    diff = np.stack(seq[:2])
    d2 = diff.copy()
    grid.kerker_prec(diff, 2, 1.0, 0.8)
    oracle.kerker_prec(d2, 2, 1.0, 0.8, gs.nr, n123, xkg)
    assert np.abs(diff - d2).max() < 1e-12 * np.abs(d2).max()


def test_anderson_pulay_broyden_over_iterations(scf, oracle):
    import pwh
    gs, rho, grid, n123, xkg = scf
    n = rho.size
    seq = _seq(rho, 7, 9)
    # Anderson: iterations 1..6 cover every branch
    st1 = [np.zeros((3, n)), np.zeros((3, n)), np.zeros((2, n))]
    st2 = [x.copy() for x in st1]
    for it, x in enumerate(seq[:6], start=1):
        a, b = x.copy(), x.copy()
        da = pwh.mix_anderson(it, 0.4, a, *st1)
        db = oracle.anderson_mix(it, 0.4, b, *st2)
        assert np.abs(a - b).max() < 1e-11 * np.abs(b).max(), it
        assert abs(da - db) <= 1e-11 * abs(db), it
        for p, q in zip(st1, st2):
            assert np.abs(p - q).max() < 1e-11 * max(np.abs(q).max(), 1e-300)
    # Pulay, mixdim 3: the warm-up branch (iter < 3) and the history branch
    md = 3
    ro1, f1 = np.zeros((md, n)), np.zeros((md, n))
    ro1[0] = rho
    ro2, f2 = ro1.copy(), f1.copy()
    nel = float(rho.sum() * gs.omega / n)
    for it, x in enumerate(seq[:6], start=1):
        a, b = x.copy(), x.copy()
        da = grid.mix_pulay(it, a, ro1, f1, md, 0.7, nel)
        db = oracle.pulay_mix(it, b, ro2, f2, md, gs.nr, n123, xkg, 0.7, gs.omega, nel)
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max(), it
        assert abs(da - db) <= 1e-11 * abs(db), it
    # Broyden, mixdim 4, iterations 1..7
    md = 4
    sb1 = [np.zeros((2, n)), np.zeros((2, n)), np.zeros((md, n)), np.zeros((md, n)), np.zeros((md, md))]
    sb1[0][:] = rho
    sb2 = [x.copy() for x in sb1]
    for it, x in enumerate(seq, start=1):
        a, b = x.copy(), x.copy()
        da = pwh.mix_broyden(it, 0.25, 0.01, a, *sb1, md)
        db = oracle.broyden_mix(it, 0.25, 0.01, b, *sb2, md)
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max(), it
        assert abs(da - db) <= 1e-11 * abs(db), it


def test_eval_enl_matches_oracle(si8, gpu_si8, oracle):
    g = oracle_globals(si8)
    rng = np.random.default_rng(10)
    nst = 12
    c0 = np.zeros((nst, g.ngwx + 3), np.complex128)
    c0[:, :si8.ngw] = si8.psi(nst, seed=11)
    focc = rng.uniform(0.0, 2.0, nst)
    enl, fnl = gpu_si8.eval_enl(c0, focc, 0.75)
    er, fr = oracle.eval_enl(g, c0, si8.ngw, focc, 0.75)
    assert abs(enl - er) < 1e-12 * abs(er)
    # rows in species, atom, lm order
    ref_rows = np.concatenate([fr[i, ia, :4] for i in range(fr.shape[0]) for ia in range(int(g.na[i]))])
    assert np.abs(fnl - ref_rows).max() < 1e-13 * np.abs(ref_rows).max()


def test_shim_eval_energy_and_mixer(si8, scf, oracle, capfd):
    """EvalEnergy and SimpleRhoMix through their mangled names with the reference globals."""
    gs, rho, _, _, _ = scf
    G = C.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=C.RTLD_GLOBAL)
    S = C.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    keep = dict(n1=np.ascontiguousarray(gs.n1, np.int32), n2=np.ascontiguousarray(gs.n2, np.int32),
                n3=np.ascontiguousarray(gs.n3, np.int32), g=np.ascontiguousarray(gs.g2),
                sfac=np.ascontiguousarray(gs.sfac.T), vps=np.ascontiguousarray(gs.vps.T),
                rhops=np.ascontiguousarray(gs.rhops.T), ngw=np.array([si8.ngw], np.int32),
                n123=np.ascontiguousarray(si8.gv["n123"]), xkg=np.ascontiguousarray(si8.gv["xkg"]))
    for name, val in (("NR1", si8.nr[0]), ("NR2", si8.nr[1]), ("NR3", si8.nr[2]), ("NNR", rho.size),
                      ("NG", len(gs.n1)), ("NSP", 1), ("NKPT", 1)):
        C.c_int.in_dll(G, name).value = int(val)
    for name, val in (("TPIBA2", gs.tpiba2), ("OMEGA", gs.omega), ("EKIN", 12.5), ("ENL", 3.25),
                      ("ESR", -1.5), ("ESELF", 7.0)):
        C.c_double.in_dll(G, name).value = float(val)
    for name, arr in (("N1", "n1"), ("N2", "n2"), ("N3", "n3"), ("G", "g"), ("SFAC", "sfac"), ("VPS", "vps"),
                      ("RHOPS", "rhops"), ("NGW", "ngw"), ("N123", "n123"), ("XKG", "xkg")):
        C.c_void_p.in_dll(G, name).value = keep[arr].ctypes.data
    f = getattr(S, "_Z10EvalEnergyPd")
    f.argtypes = [C.c_void_p]
    r = rho.copy()
    capfd.readouterr()
    f(r.ctypes.data)
    out = capfd.readouterr().out
    vr, er = oracle.eval_energy(gs, rho, ekin=12.5, enl=3.25, esr=-1.5, eself=7.0)
    assert np.abs(r - vr).max() < 1e-12 * np.abs(vr).max()
    assert abs(C.c_double.in_dll(G, "ETOT").value - er["etot"]) < 1e-12 * abs(er["etot"])
    assert abs(C.c_double.in_dll(G, "EXC").value - er["exc"]) < 1e-12 * abs(er["exc"])
    assert out.startswith("Evaluating Kohn-Sham energy terms:\n")
    assert f"ETOT   = {er['etot']:18.10f}" in out
    m = getattr(S, "_Z12SimpleRhoMixdPdS_iRd")
    m.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    a, o = rho * 1.1, rho.copy()
    b, o2 = a.copy(), o.copy()
    d = C.c_double()
    m(0.4, a.ctypes.data, o.ctypes.data, a.size, C.byref(d))
    dr = oracle.simple_mix(0.4, b, o2)
    assert np.abs(a - b).max() < 1e-15 * np.abs(b).max() and abs(d.value - dr) < 1e-14
    assert "In SimpleRhoMix: d = " in capfd.readouterr().out
    del keep
