"""GPU parity at the benchmark's larger shapes, against the FP64 oracle.

  * Si216 cell (96^3 grid, 43867 plane waves, 432 bands, 864 projectors): H|psi> of the
    whole block in FP64, FP32 and BF16x3 (the 96^3 FP64 plane pass runs one workgroup per
    CU); three bands against the oracle (1e-12 / 2e-6 / 2e-5 relative per band, as at
    Si8) and psi^H H psi Hermitian on the full block.
  * Si64 ChFSI as benchmarked (degree 10, 20-step FP64 Lanczos from a random start, 10
    outer iterations from a random orthonormal block): the Lanczos T against the oracle;
    FP32 eigenvalues within 1e-9 Ha of the FP64 run on every state FP64 has converged
    (residual < 1e-6) after 10 iterations, and on all 126 after 30; BF16x3 within 1e-7 Ha
    on the converged states (it does not reach 1e-8 Ha); orthonormal Ritz vectors and the
    residual norms of a band subset recomputed with the oracle's H|psi>
    (ChebySCFLoop.cpp:271-311,415-433).
  * 24 uneven cells (grids 24^3 to 60^3, 8 to 96 bands) through one captured filter:
    bitwise equal to the cells filtered one at a time, two cells against the oracle; the
    grouped filter (one launch per pass and kernel over all cells) bitwise equal to the
    cells filtered one at a time.
"""
import numpy as np
import pytest

from conftest import oracle_globals, relerr

pytestmark = pytest.mark.gpu


def band_relerr(a, b):
    return max(relerr(a[i], b[i]) for i in range(a.shape[0]))


@pytest.fixture(scope="module")
def si216():
    from pwh import synth
    return synth.config("si216")


def test_si216_full_block_all_paths(si216, oracle):
    import pwh
    s = si216
    assert (s.ngw, s.nbands, 4 * s.natoms) == (43867, 432, 864)
    pwh.init(0)
    g = pwh.System(s.globals())
    g.set_vloc(s.vloc)
    psi = s.psi(432, seed=31)
    idx = [0, 215, 431]
    ref = oracle.apply_ham_block(oracle_globals(s), psi[idx], s.vloc, s.ngw, nthreads=3)
    out = g.apply_ham_block(psi, "fp64")
    assert band_relerr(out[idx], ref) < 1e-12
    T = psi.conj() @ out.T
    assert np.abs(T - T.conj().T).max() < 1e-12 * np.abs(T).max()
    for prec, tol in (("fp32", 2e-6), ("bf16x3", 2e-5), ("bf16x6", 2e-6)):
        o = g.apply_ham_block(psi, prec)
        assert band_relerr(o[idx], ref) < tol, prec
        # every band against the FP64 result: the worst of 432 FP32 bands is 2.2e-6
        assert band_relerr(o, out) < 2 * tol, prec
    g.close()


@pytest.fixture(scope="module")
def si64():
    from pwh import synth
    return synth.config("si64")


@pytest.fixture(scope="module")
def si64_chfsi(si64):
    """FP64, FP32 and BF16x3 ChFSI runs from one random start (degree 10, 10 iterations,
    20-step Lanczos from a random v0)."""
    import pwh
    s = si64
    pwh.init(0)
    g = pwh.System(s.globals())
    g.set_vloc(s.vloc)
    rng = np.random.default_rng(2024)
    X0 = pwh.ortho_qr(rng.standard_normal((128, s.ngw)) + 1j * rng.standard_normal((128, s.ngw)))
    v0 = rng.random(s.ngw) + 1j * rng.random(s.ngw)
    runs = {}
    for p in ("fp64", "fp32", "bf16x3", "bf16x6"):
        runs[p] = g.chfsi(X0, degree=10, niter=10, precision=p, lanczos_k=20, v0=v0)
    for p in ("fp64", "fp32"):
        runs[p + "_30"] = g.chfsi(X0, degree=10, niter=30, precision=p, lanczos_k=20, v0=v0)
    yield g, v0, runs
    g.close()


def test_si64_lanczos_matches_oracle(si64, si64_chfsi, oracle):
    g, v0, _ = si64_chfsi
    T, f, v = g.lanczos(v0, 20)
    Tr, fr, vr = oracle.lanczos(oracle_globals(si64), v0, 20, si64.vloc, si64.ngw)
    # entries are O(10..100) Ha; 20 steps amplify 1e-16 differences in H|psi> only mildly
    assert np.abs(T - Tr).max() < 1e-9 * max(1.0, np.abs(Tr).max())
    assert relerr(v, vr) < 1e-14
    assert relerr(f, fr) < 1e-6


def test_si64_chfsi_mixed_vs_fp64(si64_chfsi):
    _, _, runs = si64_chfsi
    nx = 126
    e64, _, r64, _ = runs["fp64"]
    assert np.all(np.diff(e64) >= 0)
    # after 10 iterations the top few states are still unconverged in both runs (FP64
    # residual up to 4e-2), and their Ritz values differ by the trajectories (~2e-7 Ha);
    # on the states FP64 has converged the FP32 filter is within 1.2e-10 Ha
    conv = r64[:nx] < 1e-6
    assert conv[:90].all()
    d32 = np.abs(runs["fp32"][0][:nx] - e64[:nx])
    assert d32[conv].max() < 1e-9, d32[conv].max()
    # converged further, every state agrees (measured 1.9e-10 Ha)
    e64c, _, r64c, _ = runs["fp64_30"]
    assert r64c[:nx].max() < 1e-5
    d30 = np.abs(runs["fp32_30"][0][:nx] - e64c[:nx]).max()
    assert d30 < 1e-9, d30
    # BF16x3 keeps ~16 significant bits per GEMM product: ~3e-8 Ha on the converged
    # states, so it does not meet the 1e-8 Ha gate (FP32 does)
    dbf = np.abs(runs["bf16x3"][0][:nx] - e64[:nx])[conv].max()
    assert dbf < 1e-7, dbf
    # BF16x6 (three-term splits, six products) carries FP32-level GEMM accuracy
    d6 = np.abs(runs["bf16x6"][0][:nx] - e64[:nx])[conv].max()
    assert d6 < 1e-9, d6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_si64_chfsi_ritz_vectors_and_residuals(si64, si64_chfsi, oracle, prec):
    _, _, runs = si64_chfsi
    eig, X, res, st = runs[prec]
    assert np.abs(X.conj() @ X.T - np.eye(128)).max() < 1e-12
    idx = [0, 1, 50, 100, 125]
    HX = oracle.apply_ham_block(oracle_globals(si64), X[idx], si64.vloc, si64.ngw, nthreads=5)
    r = np.linalg.norm(HX - eig[idx, None] * X[idx], axis=1)
    # the Ritz pairs' residuals, ||H x - lambda x||, as the library reports them
    assert np.abs(res[idx] - r).max() < 1e-9 * max(1.0, r.max())
    # the lowest 90 states converge in 10 iterations (FP64 to 4e-9; the FP32 filter's
    # vectors carry its rounding, residual floor ~1.5e-4)
    assert res[:90].max() < (1e-7 if prec == "fp64" else 1e-3)


@pytest.fixture(scope="module")
def cells24():
    from pwh import synth
    return synth.uneven_batch(count=24, seed=2026)


def test_c3_24_cells_graph_equals_one_at_a_time(cells24, oracle):
    import pwh
    cells = cells24
    assert len({c.nr[0] for c in cells}) >= 6
    pwh.init(0)
    systems = []
    for c in cells:
        s = pwh.System(c.globals())
        s.set_vloc(c.vloc)
        systems.append(s)
    X0 = [c.psi(c.nbands, seed=i) for i, c in enumerate(cells)]
    lb = np.array([-1.0 + 0.05 * i for i in range(len(cells))])
    ub = np.array([45.0 + 0.5 * i for i in range(len(cells))])
    B = pwh.Batch(systems, [c.nbands for c in cells])
    for prec in ("fp64", "fp32"):
        outs = {}
        for mode in (0, 1, 1):
            for i, x in enumerate(X0):
                B.set(i, x)
            B.chebyfilt(5, lb, ub, prec, mode)
            pwh.synchronize()
            outs.setdefault(mode, []).append([B.get(i) for i in range(len(cells))])
        seq = outs[0][0]
        for rep in outs[1]:
            for i in range(len(cells)):
                assert np.array_equal(rep[i], seq[i]), (prec, i)
        if prec == "fp64":
            small = min(range(len(cells)), key=lambda i: cells[i].ngw * cells[i].nbands)
            mid = sorted(range(len(cells)), key=lambda i: cells[i].ngw)[len(cells) // 2]
            for i in (small, mid):
                c = cells[i]
                ref = oracle.chebyfilt(oracle_globals(c), X0[i], 5, lb[i], ub[i], c.vloc, c.ngw)
                assert relerr(seq[i], ref) < 1e-11, i
    B.close()
    for s in systems:
        s.close()


def test_c3_grouped_launches_equal_one_at_a_time(cells24, oracle):
    """Mode 2 (each pass of each H|psi> one launch per kernel over the 24 systems: the
    systems of one grid share a launch) gives every system its one-at-a-time filter bit for
    bit, on replays of the captured graph too, on the compiled and on the runtime FFT path;
    two cells match the oracle's chebyfilt."""
    import pwh
    cells = cells24
    pwh.init(0)
    systems = []
    for c in cells:
        s = pwh.System(c.globals())
        s.set_vloc(c.vloc)
        systems.append(s)
    X0 = [c.psi(c.nbands, seed=40 + i) for i, c in enumerate(cells)]
    lb = np.array([-0.5 + 0.03 * i for i in range(len(cells))])
    ub = np.array([40.0 + 0.7 * i for i in range(len(cells))])
    B = pwh.Batch(systems, [c.nbands for c in cells])

    def run(prec, mode, degree):
        for i, x in enumerate(X0):
            B.set(i, x)
        B.chebyfilt(degree, lb, ub, prec, mode)
        pwh.synchronize()
        return [B.get(i) for i in range(len(cells))]

    for path, cases in (("auto", (("fp64", 5), ("fp32", 5), ("fp32", 2), ("fp64", 1), ("fp32", 1))),
                        ("runtime", (("fp64", 3), ("fp32", 3)))):
        for s in systems:
            assert s.fft_path(path) in ("compiled", "runtime")
        for prec, degree in cases:
            seq = run(prec, 0, degree)
            for rep in range(2):
                grp = run(prec, 2, degree)
                for i in range(len(cells)):
                    assert np.array_equal(grp[i], seq[i]), (path, prec, degree, rep, i)
    for s in systems:
        s.fft_path("auto")
    grp = run("fp64", 2, 5)
    small = min(range(len(cells)), key=lambda i: cells[i].ngw * cells[i].nbands)
    big = max(range(len(cells)), key=lambda i: cells[i].ngw)
    for i in (small, big):
        c = cells[i]
        ref = oracle.chebyfilt(oracle_globals(c), X0[i], 5, lb[i], ub[i], c.vloc, c.ngw)
        assert relerr(grp[i], ref) < 1e-11, i
    with pytest.raises(pwh.PwhError) as e:
        B.chebyfilt(3, lb, ub, "bf16x3", 2)
    assert e.value.code == pwh.PWH_ERR_UNSUPPORTED
    with pytest.raises(pwh.PwhError) as e:
        B.chebyfilt(3, lb, ub, "fp32", 3)
    assert e.value.code == pwh.PWH_ERR_ARG
    B.close()
    for s in systems:
        s.close()


def test_si216_grouped_filter_equals_one_at_a_time(si216):
    """The 96^3 grid in a grouped batch (mode 2): two copies of the Si216 system with 16
    states each, FP32 filter (the 96 x 96 four-step plane kernel, grouped form) and FP64
    (the Stockham kernel), every system bit for bit equal to its one-at-a-time filter."""
    import pwh
    s = si216
    pwh.init(0)
    systems = []
    for _ in range(2):
        g = pwh.System(s.globals())
        g.set_vloc(s.vloc)
        systems.append(g)
    X0 = [s.psi(16, seed=70 + i) for i in range(2)]
    lb, ub = np.array([-0.5, -0.4]), np.array([40.0, 41.0])
    B = pwh.Batch(systems, [16, 16])

    def run(prec, mode):
        for i, x in enumerate(X0):
            B.set(i, x)
        B.chebyfilt(3, lb, ub, prec, mode)
        pwh.synchronize()
        return [B.get(i) for i in range(2)]

    for prec in ("fp32", "fp64"):
        seq = run(prec, 0)
        grp = run(prec, 2)
        for i in range(2):
            assert np.array_equal(grp[i], seq[i]), (prec, i)
    B.close()
    for g in systems:
        g.close()
