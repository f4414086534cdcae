"""Pins the CPU oracle (oracle/oracle.c + oracle/oracle.py) before it is trusted.

Each restated function is checked against something other than its own formula:
the reference's compiled code (oracle/_ref, when built), the known answers printed in
the reference's logs, committed golden fixtures written from the reference, numpy,
brute-force sums, and closed forms.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_globals, relerr

GOLDEN = os.path.join(ROOT, "tests", "golden")


def have_ref():
    from oracle import oracle as O
    return O.Reference.available()


# --------------------------------------------------------------------------
# FFT conventions (fft_fftw3.cpp:13-46)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nr", [(3, 3, 2), (6, 8, 10), (12, 10, 14), (32, 32, 32), (5, 7, 9),
                                (22, 12, 16), (14, 26, 38), (82, 6, 4), (41, 43, 2), (28, 28, 28),
                                (70, 4, 140)])
def test_fft3d_matches_numpy(oracle, nr):
    rng = np.random.default_rng(0)
    n = int(np.prod(nr))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = x.reshape(nr[::-1])                       # [i3][i2][i1], x fastest
    inv = oracle.fft3d(x, nr, True)
    fwd = oracle.fft3d(x, nr, False)
    assert relerr(inv, np.fft.ifftn(a).reshape(-1) * n) < 1e-14    # exp(+i), unscaled
    assert relerr(fwd, np.fft.fftn(a).reshape(-1) / n) < 1e-14     # exp(-i), 1/N


def test_fft3d_brute_force_dft(oracle):
    nr = (3, 4, 5)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    i1, i2, i3 = np.meshgrid(np.arange(3), np.arange(4), np.arange(5), indexing="ij")
    r = np.stack([i1.ravel(), i2.ravel(), i3.ravel()], 1)             # grid points
    idx = r[:, 0] + 3 * (r[:, 1] + 4 * r[:, 2])
    phase = (np.outer(r[:, 0], r[:, 0]) / 3 + np.outer(r[:, 1], r[:, 1]) / 4
             + np.outer(r[:, 2], r[:, 2]) / 5)
    D = np.exp(2j * np.pi * phase)                                    # D[k, j]
    xs = x[idx]
    want = np.zeros(60, complex)
    want[idx] = D @ xs
    assert relerr(oracle.fft3d(x, nr, True), want) < 1e-13


def test_fft3d_sandbox_known_answer(oracle):
    # sandbox/t_fft_fftw3.cpp:18-31: 3x3x2 grid of (1,2), inverse transform
    out = oracle.fft3d(np.full(18, 1 + 2j), (3, 3, 2), True)
    assert abs(out[0] - 18 * (1 + 2j)) < 1e-13
    assert np.abs(out[1:]).max() < 1e-13


# --------------------------------------------------------------------------
# G-vector maps and phases
# --------------------------------------------------------------------------

# (lattice a [bohr], ecut [Ry], grid) -> (NGW, NG, NGWX, NGX, GCUT) printed by the reference
LOG_KNOWN = [
    (10.266, 8.0, (20, 20, 20), 437, 3287, 446, 3568, 85.4265291430),     # works/Si8/log_8Ry_cheby:54-62
    (10.266, 15.0, (28, 28, 28), 1045, 8409, 1060, 8480, 160.1747421432),  # works/Si8/log_15Ry_lobpcg:53-62
    (20.532, 8.0, (38, 38, 38), 3287, 26529, 3325, 26600, 341.7061165722),  # works/Si64/log_davidson_8Ry:107-116
    (20.532, 10.0, (48, 48, 48), 4625, 37121, 4649, 37192, 427.1326457152),  # works/Si64/log_lobpcg_10Ry_SCF_anderson_0.3:107-116
]


@pytest.mark.parametrize("a,ecut,nr,ngw,ng,ngwx,ngx,gcut", LOG_KNOWN)
def test_gvector_counts_match_reference_logs(oracle, a, ecut, nr, ngw, ng, ngwx, ngx, gcut):
    A = np.diag([a, a, a])
    gv = oracle.gvectors(A, ecut, nr)
    assert (gv["ngw"][0], gv["ng"], gv["ngwx"], gv["ngx"]) == (ngw, ng, ngwx, ngx)
    assert abs(oracle.setup_scalars(A, ecut)["gcut"] - gcut) < 1e-9


def test_fft_grid_matches_logs(oracle):
    # works/Si8/log_15Ry_lobpcg:5,54 (minimum 25 -> 28) and log_8Ry_cheby (20), Si64 8 Ry (38)
    assert oracle.fft_grid(np.diag([10.266] * 3), 15.0) == (28, 28, 28)
    assert oracle.fft_grid(np.diag([10.266] * 3), 8.0) == (20, 20, 20)
    assert oracle.fft_grid(np.diag([20.532] * 3), 8.0) == (38, 38, 38)


def test_reference_example_grids():
    """Every FFT grid the reference's example inputs lead to (tests/golden/ref_example_grids.json:
    Setup.cpp:20-34's rule, a user fft_grid, or a grid its logs print) is one the library
    transforms, and the library's Setup restatement reproduces the rule's grids."""
    import json
    import os
    import pwh
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_example_grids.json")))
    assert len(d) >= 30
    for name, c in d.items():
        assert pwh.grid_supported(c["grid"]), (name, c["grid"])
        if "setup_grid" in c:
            assert tuple(pwh.fft_grid(np.array(c["A"]).reshape(3, 3), c["ecut"])) == tuple(c["setup_grid"]), name


def test_gvector_invariants(oracle):
    A = np.diag([10.26] * 3)
    nr = (32, 32, 32)
    gv = oracle.gvectors(A, 15.0, nr)
    ng, ngx = gv["ng"], gv["ngx"]
    IN = gv["in123"].reshape(3, ngx)[:, :ng].T
    assert np.all(IN[0] == 0)                                   # G = 0 first
    assert np.all(IN[1::2][: (ng - 1) // 2] == -IN[2::2][: (ng - 1) // 2])   # (+G, -G) pairs
    b = 1.0                                                     # cubic: |G|^2 = |n|^2 (2pi/a)^2
    g2 = (IN ** 2).sum(1) * b
    assert np.all(np.diff(g2[1::2]) >= 0)                       # sorted shells
    n = gv["ngw"][0]
    n123 = gv["n123"][:n]
    assert len(np.unique(n123)) == n and n123.min() >= 1 and n123.max() <= np.prod(nr)
    xkg = gv["xkg"][:n]
    assert np.all(xkg <= oracle.setup_scalars(A, 15.0)["gcutw"])
    # N123 decodes to the signed Miller index of IGK
    igk = gv["igk"][:n] - 1
    m = IN[igk]
    dec = np.stack([(n123 - 1) % 32, ((n123 - 1) // 32) % 32, (n123 - 1) // 1024], 1)
    assert np.array_equal(np.mod(m, 32), dec)
    assert np.allclose(xkg, (m ** 2).sum(1), rtol=0, atol=1e-12)


def test_sort1_matches_heap_semantics(oracle):
    rng = np.random.default_rng(2)
    arr = rng.integers(0, 20, 300).astype(float)                # many ties
    idx = oracle.sort1(arr)
    assert sorted(idx) == list(range(1, 301))
    assert np.all(np.diff(arr[idx - 1]) >= 0)
    if have_ref():
        from oracle.oracle import Reference
        assert np.array_equal(idx, Reference().sort1(arr))


def _diamond(a, ncell):
    from pwh import synth
    return synth.diamond_fractional(ncell) * a


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("a,ecut,ncell,nr", [(10.266, 8.0, 1, (20, 20, 20)), (10.26, 15.0, 1, (32, 32, 32)),
                                             (10.0, 12.0, 1, (24, 30, 36)), (20.52, 20.0, 2, (64, 64, 64)),
                                             (7.5, 9.0, 1, (16, 20, 18))])
def test_oracle_maps_equal_compiled_reference(oracle, a, ecut, ncell, nr):
    from oracle.oracle import Reference
    A = np.diag([a, a, a])
    tau = _diamond(a, ncell)
    nat = len(tau)
    gv = oracle.gvectors(A, ecut, nr)
    r = Reference().setup(A, ecut, nr, [nat], nat, tau[None])
    for k in ("in123", "igk", "xkg", "n123", "ggk"):
        assert np.array_equal(gv[k], r[k]), k
    assert gv["ng"] == r["ng"] and gv["ngwx"] == r["ngwx"] and gv["ngx"] == r["ngx"]
    eigr = oracle.phase_factor(A, nr, [nat], nat, tau[None], gv)
    assert np.array_equal(eigr, r["eigr"])


def test_oracle_maps_equal_golden_fixtures(oracle):
    """Fixtures written from the compiled reference by tests/golden/make_golden.py."""
    files = sorted(f for f in os.listdir(GOLDEN) if f.startswith("ref_maps_") and f.endswith(".npz"))
    assert files, "golden fixtures missing"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        A = z["A"]
        gv = oracle.gvectors(A, float(z["ecut"]), tuple(z["nr"]))
        n = int(z["ngw"])
        assert gv["ngw"][0] == n
        assert np.array_equal(gv["n123"][:n], z["n123"]), f
        assert np.array_equal(gv["igk"][:n], z["igk"]), f
        assert np.array_equal(gv["xkg"][:n], z["xkg"]), f
        tau = z["tau"]
        nat = tau.shape[0]
        eigr = oracle.phase_factor(A, tuple(z["nr"]), [nat], nat, tau[None], gv)
        e = eigr.reshape(nat, gv["ngwx"])[:, :n]
        assert np.array_equal(e, z["eigr"]), f


def test_eigr_is_structure_factor(oracle):
    a = 10.26
    A = np.diag([a] * 3)
    nr = (32, 32, 32)
    tau = _diamond(a, 1)
    gv = oracle.gvectors(A, 15.0, nr)
    n = gv["ngw"][0]
    e = oracle.phase_factor(A, nr, [8], 8, tau[None], gv).reshape(8, gv["ngwx"])[:, :n]
    IN = gv["in123"].reshape(3, gv["ngx"]).T[gv["igk"][:n] - 1]
    G = IN * (2 * np.pi / a)
    want = np.exp(-1j * tau @ G.T)
    assert np.abs(e - want).max() < 1e-13


# --------------------------------------------------------------------------
# H|psi> (ApplyHam_block.cpp:5-172)
# --------------------------------------------------------------------------

def small_system(lm_end=4, ecut=9.0, a=7.0, nr=(12, 10, 14), natoms=2):
    from pwh import synth
    return synth.make_system(a, ecut, 1, nr, 4, name="small", natoms=natoms)


def globals_of(s, lmax=3, lloc=3, vloc=None):
    from conftest import oracle_globals
    g = oracle_globals(s)
    g.l_max = np.array([lmax], np.int32)
    g.l_loc = np.array([lloc], np.int32)
    return g


def test_hpsi_constant_potential_closed_form(oracle):
    s = small_system()
    g = globals_of(s, lmax=1, lloc=1)                    # lm_end = 0: no nonlocal part
    V0 = -0.731
    psi = s.psi(3, seed=3)
    out = oracle.apply_ham_block(g, psi, np.full(s.vloc.size, V0), s.ngw)
    kin = 0.5 * s.gv["xkg"][: s.ngw] * s.tpiba2
    assert relerr(out, (kin + V0)[None, :] * psi) < 1e-14


def test_hpsi_local_part_brute_force(oracle):
    """V part against explicit plane-wave matrix elements V(G - G') from a direct DFT sum."""
    s = small_system()
    g = globals_of(s, lmax=1, lloc=1)
    nr = s.nr
    rng = np.random.default_rng(4)
    V = rng.standard_normal(s.vloc.size)
    psi = s.psi(2, seed=5)
    out = oracle.apply_ham_block(g, psi, V, s.ngw)
    kin = 0.5 * s.gv["xkg"][: s.ngw] * s.tpiba2
    n123 = s.gv["n123"][: s.ngw].astype(np.int64) - 1
    m = np.stack([n123 % nr[0], (n123 // nr[0]) % nr[1], n123 // (nr[0] * nr[1])], 1)
    i1, i2, i3 = np.meshgrid(np.arange(nr[0]), np.arange(nr[1]), np.arange(nr[2]), indexing="ij")
    r = np.stack([i1.ravel(), i2.ravel(), i3.ravel()], 1)
    Vr = V[r[:, 0] + nr[0] * (r[:, 1] + nr[1] * r[:, 2])]
    # V(G - G') = (1/N) sum_r V(r) exp(-2 pi i (G - G').r / N_d)
    dm = m[:, None, :] - m[None, :, :]
    ph = (dm[..., 0, None] * r[None, None, :, 0] / nr[0] + dm[..., 1, None] * r[None, None, :, 1] / nr[1]
          + dm[..., 2, None] * r[None, None, :, 2] / nr[2])
    Vg = (np.exp(-2j * np.pi * ph) @ Vr) / np.prod(nr)
    want = psi @ Vg.T + kin[None, :] * psi
    assert relerr(out, want) < 1e-12


def test_hpsi_nonlocal_dense_beta(oracle):
    s = small_system()
    g = globals_of(s)                                     # lm_end = 4
    psi = s.psi(3, seed=6)
    out = oracle.apply_ham_block(g, psi, np.zeros(s.vloc.size), s.ngw)
    n = s.ngw
    kin = 0.5 * s.gv["xkg"][:n] * s.tpiba2
    nat = s.natoms
    E = s.eigr.reshape(nat, s.gv["ngwx"])[:, :n]          # EIGR(G, a)
    P = s.pkg_a.reshape(s.gv["ngwx"], 4)[:n].T             # PKG_A(l, G)
    beta = (E[:, None, :] * P[None, :, :]).reshape(nat * 4, n)
    D = np.tile(s.wnl, nat)
    want = kin[None, :] * psi + ((psi @ beta.conj().T) * D[None, :]) @ beta
    assert relerr(out, want) < 1e-12


def test_hpsi_hermitian(oracle):
    s = small_system()
    g = globals_of(s)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((2, s.ngw)) + 1j * rng.standard_normal((2, s.ngw))
    Ha = oracle.apply_ham_block(g, a, s.vloc, s.ngw)
    lhs, rhs = np.vdot(a[0], Ha[1]), np.conj(np.vdot(a[1], Ha[0]))
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


def test_hpsi_abort_conditions(oracle):
    s = small_system()
    g = globals_of(s)
    with pytest.raises(RuntimeError):
        oracle.apply_ham_block(g, s.psi(1), s.vloc[:-1], s.ngw)          # nnr != NNR
    g = globals_of(s, lmax=4, lloc=3)                                     # lm_end = 11 > 4
    with pytest.raises(RuntimeError):
        oracle.apply_ham_block(g, s.psi(1), s.vloc, s.ngw)


# --------------------------------------------------------------------------
# chebyfilt, lanczos, bounds
# --------------------------------------------------------------------------

def cheb_T(m, x):
    x = np.asarray(x, float)
    out = np.empty_like(x)
    big = np.abs(x) > 1
    out[~big] = np.cos(m * np.arccos(x[~big]))
    out[big] = np.sign(x[big]) ** m * np.cosh(m * np.arccosh(np.abs(x[big])))
    return out


@pytest.mark.parametrize("degree", [1, 2, 5, 10])
def test_chebyfilt_closed_form_diagonal_h(oracle, degree):
    s = small_system()
    g = globals_of(s, lmax=1, lloc=1)
    V0 = 0.25
    lam = 0.5 * s.gv["xkg"][: s.ngw] * s.tpiba2 + V0
    lb, ub = 0.9, float(lam.max()) + 0.1
    X = s.psi(3, seed=8)
    out = oracle.chebyfilt(g, X, degree, lb, ub, np.full(s.vloc.size, V0), s.ngw)
    e, c = (ub - lb) / 2, (ub + lb) / 2
    scale = cheb_T(degree, (lam - c) / e) / cheb_T(degree, np.array([-2.0]))[0]
    assert relerr(out, X * scale[None, :]) < 1e-12


def test_lanczos_relation(oracle):
    s = small_system()
    g = globals_of(s)
    rng = np.random.default_rng(9)
    v0 = rng.random(s.ngw) + 1j * rng.random(s.ngw)
    k = 15
    T, f, v, V = oracle.lanczos(g, v0, k, s.vloc, s.ngw, want_v=True)
    assert abs(np.linalg.norm(v) - 1) < 1e-14
    assert np.abs(V.conj() @ V.T - np.eye(k)).max() < 1e-12
    Tu = np.triu(T) + np.triu(T, 1).T                      # what dsyev('U') sees
    HV = oracle.apply_ham_block(g, V, s.vloc, s.ngw)
    R = HV - Tu.T @ V                                       # rows: H v_j - sum_i T_ij v_i
    R[-1] -= f
    assert np.abs(R).max() < 1e-10 * np.abs(HV).max()
    # subdiagonal holds ||f_j|| = T(j-1, j) up to the re-orthogonalisation terms
    assert np.allclose(np.diag(T, -1), np.diag(T, 1), rtol=1e-10, atol=1e-12)


def test_induced_norm_mat_is_squared_norm(oracle):
    rng = np.random.default_rng(10)
    T = rng.standard_normal((12, 12))
    assert abs(oracle.induced_norm_mat(T, 10) - np.linalg.norm(T[:10, :10] @ np.full(10, 10 ** -0.5)) ** 2) < 1e-12


# --------------------------------------------------------------------------
# dense steps vs the reference's LAPACK wrappers
# --------------------------------------------------------------------------

def test_ortho_qr_and_eig_against_golden(oracle):
    z = np.load(os.path.join(GOLDEN, "ref_lapack.npz"))
    Q = oracle.ortho_qr(z["X"])
    assert np.abs(Q - z["Q"]).max() < 1e-12
    w, _ = oracle.eig_zheevd(z["T"])
    assert np.abs(w - z["w"]).max() < 1e-12
    # sandbox/t_zheevd.cpp:41-69 4x4 Hermitian matrix
    w4, _ = oracle.eig_zheevd(z["A4"])
    assert np.abs(w4 - z["w4"]).max() < 1e-12


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference tree absent)")
def test_ortho_qr_and_eig_against_compiled_reference(oracle):
    from oracle.oracle import Reference
    R = Reference()
    rng = np.random.default_rng(11)
    X = rng.standard_normal((6, 40)) + 1j * rng.standard_normal((6, 40))
    assert np.abs(oracle.ortho_qr(X) - R.ortho_qr(X)).max() < 1e-12
    H = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    H = H + H.conj().T
    assert np.abs(oracle.eig_zheevd(H)[0] - R.eig_zheevd(H)[0]).max() < 1e-12


# ---- EvalRhoPsi (EvalRhoPsi.cpp:83-157) ------------------------------------------------

def _signed_index(n123, nr):
    """1-based N123 -> signed grid indices (n1, n2, n3) and the 0-based grid position."""
    idx = int(n123) - 1
    i1, i2, i3 = idx % nr[0], (idx // nr[0]) % nr[1], idx // (nr[0] * nr[1])
    s = [i if i < n // 2 else i - n for i, n in zip((i1, i2, i3), nr)]
    return s, idx


def test_rho_single_plane_wave_is_constant(oracle, si8):
    g = oracle_globals(si8)
    n = si8.ngw
    ev = np.zeros((2, n), np.complex128)
    ev[0, 5] = np.exp(0.3j)
    ev[1, 17] = 1.0
    w, omega = np.array([2.0, 0.5]), 1080.0
    rho, ek, charge, bad = oracle.eval_rho_psi(g, ev, n, w, omega)
    assert bad == -1 and np.allclose(charge, 1.0)
    assert np.abs(rho - (2.0 + 0.5) / omega).max() < 1e-15
    xkg = si8.gv["xkg"]
    assert abs(ek - (2.0 * 0.5 * xkg[5] + 0.5 * 0.5 * xkg[17]) * g.tpiba2) < 1e-13


def test_rho_two_plane_waves_closed_form(oracle, si8):
    g = oracle_globals(si8)
    n, nr = si8.ngw, si8.nr
    a, b, c1, c2 = 3, 40, 0.6 + 0.2j, -0.3 + 0.7j
    norm = np.sqrt(abs(c1) ** 2 + abs(c2) ** 2)
    c1, c2 = c1 / norm, c2 / norm
    ev = np.zeros((1, n), np.complex128)
    ev[0, a], ev[0, b] = c1, c2
    rho, _, charge, bad = oracle.eval_rho_psi(g, ev, n, [1.0], 1.0)
    (ga, _), (gb, _) = _signed_index(si8.gv["n123"][a], nr), _signed_index(si8.gv["n123"][b], nr)
    i3, i2, i1 = np.meshgrid(np.arange(nr[2]), np.arange(nr[1]), np.arange(nr[0]), indexing="ij")
    ph = 2 * np.pi * ((ga[0] - gb[0]) * i1 / nr[0] + (ga[1] - gb[1]) * i2 / nr[1] + (ga[2] - gb[2]) * i3 / nr[2])
    ref = abs(c1) ** 2 + abs(c2) ** 2 + 2 * np.real(c1 * np.conj(c2) * np.exp(1j * ph))
    assert bad == -1 and abs(charge[0] - 1.0) < 1e-14
    assert np.abs(rho - ref.reshape(-1)).max() < 1e-13


def test_rho_parseval_and_accumulation(oracle, si8):
    g = oracle_globals(si8)
    n = si8.ngw
    X = si8.psi(6, seed=41)
    X[:, n - 1] = 0.0          # the reference never reads the last coefficient
    X /= np.linalg.norm(X, axis=1)[:, None]
    w = np.array([2.0, 2.0, 2.0, 1.0, 0.5, 0.0])
    base = np.full(int(np.prod(si8.nr)), 0.25)
    rho, _, charge, bad = oracle.eval_rho_psi(g, X, n, w, 300.0, rhoe=base)
    assert bad == -1
    assert abs((rho - base).sum() - np.prod(si8.nr) * w.sum() / 300.0) < 1e-9
    assert np.all(rho - base >= -1e-15)


def test_rho_skips_last_coefficient_and_checks_charge(oracle, si8):
    g = oracle_globals(si8)
    n = si8.ngw
    ev = np.zeros((3, n), np.complex128)
    ev[0, 2] = np.sqrt(1.005)          # inside the 6e-3 window
    ev[1, 9] = np.sqrt(1.007)          # outside: the reference aborts here
    ev[2, n - 1] = 1.0                 # only the skipped coefficient
    rho, _, charge, bad = oracle.eval_rho_psi(g, ev, n, [1.0, 1.0, 1.0], 1.0)
    assert bad == 1 and abs(charge[1] - 1.007) < 1e-12
    assert np.abs(rho - 1.005).max() < 1e-13     # state 0's density only
    _, _, charge2, bad2 = oracle.eval_rho_psi(g, ev[2:], n, [1.0], 1.0)
    assert bad2 == 0 and charge2[0] == 0.0
