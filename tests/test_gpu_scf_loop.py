"""The whole Kohn-Sham cycle on the GPU (pwh.scf.cheby_scf, ChebySCFLoop.cpp:171-372 over
the library's entry points) on the Si8 cell: ChFSI step with V_eff, EvalRhoPsi, EvalENL,
mixing, EvalEnergy.

* FP64 trajectory against the oracle composed the same way on the CPU (oracle chebyfilt,
  ortho_qr, apply_ham_block + Rayleigh-Ritz, eval_rho_psi, eval_enl, simple_mix,
  eval_energy): ETOT of every iteration to 1e-8 relative.
* Mixed precision against FP64 at self-consistency: with an FP32 filter throughout, the
  converged energy wanders with the filter's rounding, its median distance to the FP64
  energy below 1e-6 Ha (the acceptance figure) and the occupied eigenvalues within 5e-8 Ha;
  with the last iterations in FP64 (fp64_refine), the FP64 fixed point to 1e-8 Ha.

The synthetic cell's levels come in degenerate shells (6, 6, 3, 3, 6 states at the bottom);
12 occupied states close a shell, so the loop converges with the reference's integer
occupations (16 would half-fill a shell, which needs smearing the reference lacks), and 15
states close the next one, so the subspace edge does not split a multiplet."""
import numpy as np
import pytest

from conftest import density_gspace, oracle_globals

pytestmark = pytest.mark.gpu

NOCC, NST, DEGREE, CHFSI_ITERS, BETA = 12, 15, 10, 2, 0.3


@pytest.fixture(scope="module")
def scf_setup(si8):
    import pwh
    from pwh import synth
    pwh.init(0)
    d = synth.density_gspace(si8)
    grid = pwh.Grid(si8.nr, d["n1"], d["n2"], d["n3"], d["g2"], d["sfac"], d["vps"], d["rhops"], d["tpiba2"],
                    d["omega"], n123=si8.gv["n123"][:si8.ngw].astype(np.int64), xkg=si8.gv["xkg"][:si8.ngw])
    sysm = pwh.System(si8.globals())
    rng = np.random.default_rng(5)
    X0 = pwh.ortho_qr(rng.standard_normal((NST, si8.ngw)) + 1j * rng.standard_normal((NST, si8.ngw)))
    rho0 = np.full(int(np.prod(si8.nr)), 2.0 * NOCC / d["omega"])
    yield sysm, grid, X0, rho0, d["omega"]
    sysm.close()
    grid.close()


def _v0(ngw):
    rng = np.random.default_rng(17)
    return rng.random(ngw) + 1j * rng.random(ngw)


def _run(scf_setup, precision, max_iter=200, conv=1e-10, **kw):
    from pwh.scf import cheby_scf
    sysm, grid, X0, rho0, omega = scf_setup
    return cheby_scf(sysm, grid, X0, NOCC, np.full(NOCC, 2.0), omega, rho0, degree=DEGREE, precision=precision,
                     mixer="simple", beta=BETA, max_iter=max_iter, conv=conv, chfsi_iters=CHFSI_ITERS,
                     v0=_v0(X0.shape[1]), **kw)


def _oracle_scf(oracle, si8, X0, rho0, omega, iters):
    """The same cycle from the oracle's pieces (bounds: 20-step Lanczos once, ub = largest
    Ritz value + ||f||, lb = w0 + frac (ub - w0); after each Rayleigh-Ritz step lb = its
    largest value)."""
    g = oracle_globals(si8)
    gs = density_gspace(si8)
    ngw = si8.ngw
    rho = rho0.copy()
    rho_old = rho.copy()
    veff, _ = oracle.eval_energy(gs, rho.copy(), 0.0, 0.0, 0.0, 0.0)
    T, f, _ = oracle.lanczos(g, _v0(ngw), 20, veff, ngw)
    w = np.linalg.eigvalsh(T, UPLO="U")    # dsyev uplo U, as ChebySCFLoop.cpp:180
    ub = w[-1] + np.sqrt(np.vdot(f, f).real)
    frac = min(0.5, max(0.02, (NST / ngw) ** (2.0 / 3.0)))
    lb = w[0] + frac * (ub - w[0])
    X = X0.copy()
    out = []
    for _ in range(iters):
        for _ in range(CHFSI_ITERS):
            X = oracle.chebyfilt(g, X, DEGREE, lb, ub, veff, ngw)
            X = oracle.ortho_qr(X)
            HX = oracle.apply_ham_block(g, X, veff, ngw)
            eig, X = oracle.rayleigh_ritz(X, HX)
            lb = eig[-1]
        focc = np.full(NOCC, 2.0)
        rho, ekin, _, bad = oracle.eval_rho_psi(g, X[:NOCC], ngw, focc, omega)
        assert bad < 0
        c0 = np.zeros((NOCC, g.ngwx + 3), np.complex128)
        c0[:, :ngw] = X[:NOCC]
        enl, _ = oracle.eval_enl(g, c0, ngw, focc, 1.0)
        oracle.simple_mix(BETA, rho, rho_old)
        veff, e = oracle.eval_energy(gs, rho.copy(), ekin, enl, 0.0, 0.0)
        out.append(e["etot"])
    return out


def test_scf_fp64_trajectory_matches_oracle(scf_setup, si8, oracle):
    sysm, grid, X0, rho0, omega = scf_setup
    r = _run(scf_setup, "fp64", max_iter=3, conv=0.0)
    ref = _oracle_scf(oracle, si8, X0, rho0, omega, 3)
    got = [h[1] for h in r.history]
    assert len(got) == 3
    assert np.abs(np.array(got) - np.array(ref)).max() < 1e-8 * abs(ref[-1]), (got, ref)


def test_scf_converges_and_mixed_total_energy_within_1e6(scf_setup):
    import pwh
    r64 = _run(scf_setup, "fp64")
    assert r64.converged and r64.iterations < 200
    # the energy terms add up (EvalEnergy.cpp:142-144; EHT carries ESR - ESELF)
    e = r64.energies
    parts = e["ekin"] + e["enl"] + e["eht"] + e["epseu"] + e["exc"]
    assert abs(e["etot"] - parts) < 1e-9 * abs(e["etot"])
    # mixed filters throughout: past convergence the energy wanders with the filter's
    # rounding; its median distance to the FP64 fixed point is the figure
    for prec, tol in (("fp32", 1e-6), ("bf16x3", 1e-4)):
        r = _run(scf_setup, prec, max_iter=150, conv=-1.0)
        dev = np.abs(np.array([h[1] for h in r.history[100:]]) - r64.etot)
        assert np.median(dev) < tol, (prec, np.median(dev), dev.max())
        if prec == "fp32":
            assert np.abs(r.eig[:NOCC] - r64.eig[:NOCC]).max() < 5e-8
    # FP32 until |dE| <= 1e-6, then FP64: the FP64 fixed point
    r = _run(scf_setup, "fp32", fp64_refine=1e-6)
    assert r.converged
    n32 = sum(1 for h in r.history if h[6] == pwh.FP32)
    assert 10 < n32 < r.iterations
    assert abs(r.etot - r64.etot) < 1e-8, (r.etot, r64.etot)
    assert np.abs(r.eig[:NOCC] - r64.eig[:NOCC]).max() < 1e-9


def test_scf_resident_and_host_paths_agree(scf_setup):
    """The resident entry points (block kept in HBM) and the host-block ones give the same
    steps: ETOT per step to 1e-10 relative, final eigenvalues to 1e-10 Ha."""
    a = _run(scf_setup, "fp64", max_iter=6, conv=0.0, resident=True)
    b = _run(scf_setup, "fp64", max_iter=6, conv=0.0, resident=False)
    ea, eb = np.array([h[1] for h in a.history]), np.array([h[1] for h in b.history])
    assert np.abs(ea - eb).max() < 1e-10 * abs(eb[-1])
    assert np.abs(a.eig - b.eig).max() < 1e-10


@pytest.mark.parametrize("mixer", ["simple", "kerker", "anderson", "broyden"])
def test_scf_device_state_matches_host_arrays(scf_setup, mixer):
    """Density, V_eff and mixing history held in HBM (pwh_scf_state_*) against the same loop
    with the host-array entry points: the same kernels in the same order, so ETOT of every
    step, the eigenvalues and the final density are equal bit for bit."""
    kw = dict(max_iter=5, conv=0.0, resident=True)
    from pwh.scf import cheby_scf
    sysm, grid, X0, rho0, omega = scf_setup
    runs = []
    for dev in (True, False):
        runs.append(cheby_scf(sysm, grid, X0, NOCC, np.full(NOCC, 2.0), omega, rho0, degree=DEGREE,
                              precision="fp64", mixer=mixer, beta=BETA, chfsi_iters=CHFSI_ITERS,
                              v0=_v0(X0.shape[1]), device_state=dev, **kw))
    a, b = runs
    ea, eb = np.array([h[1] for h in a.history]), np.array([h[1] for h in b.history])
    assert np.array_equal(ea, eb), (mixer, ea - eb)
    assert np.array_equal([h[3] for h in a.history], [h[3] for h in b.history])
    assert np.array_equal(a.eig, b.eig)
    assert np.array_equal(a.rho, b.rho)


def test_scf_state_steps_match_host_entry_points(scf_setup):
    """Each pwh_scf_state step against its host-array entry point on the same inputs:
    EvalEnergy (energies and V_eff), the local potential it sets (H|psi> unchanged) and the
    density of the resident states (the mixing steps are checked through the loop above).
    Pulay is not among the state's mixers."""
    import pwh
    sysm, grid, X0, rho0, omega = scf_setup
    focc = np.full(NOCC, 2.0)
    rng = np.random.default_rng(3)
    rho = rho0 * (1.0 + 0.1 * rng.random(rho0.size))
    st = pwh.ScfState(grid, "anderson")
    try:
        st.set_rho(rho)
        e1 = st.energy(1.5, 0.25, 0.1, 0.05)
        v2, e2 = grid.eval_energy(rho, 1.5, 0.25, 0.1, 0.05)
        r1, v1 = st.get()
        assert np.array_equal(r1, rho) and np.array_equal(v1, v2) and e1 == e2
        st.set_vloc(sysm)
        h1 = sysm.apply_ham_block(X0[:4], "fp64")
        h1s = sysm.apply_ham_block(X0[:4], "fp32")
        sysm.set_vloc(v2)
        assert np.array_equal(h1, sysm.apply_ham_block(X0[:4], "fp64"))
        assert np.array_equal(h1s, sysm.apply_ham_block(X0[:4], "fp32"))
        sysm.resident_set(X0)
        k1, c1 = st.density(sysm, NOCC, focc, omega)
        r2, k2, c2 = sysm.resident_eval_rho_psi(NOCC, focc, omega)
        assert np.array_equal(st.get()[0], r2) and k1 == k2 and np.array_equal(c1, c2)
        with pytest.raises(pwh.PwhError):
            st.density(sysm, NST + 1, np.full(NST + 1, 2.0), omega)
    finally:
        st.close()
    with pytest.raises(pwh.PwhError):
        pwh.ScfState(grid, 3)


def test_scf_autotuned_targets_keep_the_trajectory(scf_setup):
    """With the split-K targets tuned first the steps only change by the order of the
    split-K sums: ETOT per step to 1e-10 relative."""
    a = _run(scf_setup, "fp64", max_iter=4, conv=0.0)
    b = _run(scf_setup, "fp64", max_iter=4, conv=0.0, autotune=True)
    ea, eb = np.array([h[1] for h in a.history]), np.array([h[1] for h in b.history])
    assert np.abs(ea - eb).max() < 1e-10 * abs(ea[-1])
    with pytest.raises(ValueError):
        _run(scf_setup, "fp64", max_iter=1, autotune=True, resident=False)


def test_resident_entry_points_match_host_ones(scf_setup):
    import pwh
    sysm, grid, X0, rho0, omega = scf_setup
    focc = np.full(NOCC, 2.0)
    sysm.resident_set(X0)
    before = sysm.resident_get(NST)
    r1, k1, c1 = sysm.resident_eval_rho_psi(NOCC, focc, omega)
    e1, f1 = sysm.resident_eval_enl(NOCC, focc, 0.5)
    # the resident block is unchanged by both (the density works on a copy)
    assert np.array_equal(sysm.resident_get(NST), before)
    r2, k2, c2 = sysm.eval_rho_psi(X0[:NOCC], focc, omega)
    e2, f2 = sysm.eval_enl(X0[:NOCC], focc, 0.5)
    assert np.array_equal(r1, r2) and k1 == k2 and np.array_equal(c1, c2)
    assert e1 == e2 and np.array_equal(f1, f2)
    sysm.resident_set(X0)
    Xr = sysm.resident_get(NST)
    assert np.abs(Xr - X0).max() == 0.0
    with pytest.raises(pwh.PwhError):
        sysm.resident_eval_rho_psi(NST + 1, np.full(NST + 1, 2.0), omega)


def test_scf_driver_rejects_bad_inputs(scf_setup):
    from pwh.scf import cheby_scf
    sysm, grid, X0, rho0, omega = scf_setup
    with pytest.raises(ValueError):
        cheby_scf(sysm, grid, X0, NOCC, np.full(NOCC, 2.0), omega, rho0[:-1])
    with pytest.raises(ValueError):
        cheby_scf(sysm, grid, X0, NOCC, np.full(NOCC - 1, 2.0), omega, rho0)
    with pytest.raises(ValueError):
        cheby_scf(sysm, grid, X0, NOCC, np.full(NOCC, 2.0), omega, rho0, mixer="pulay")
