"""Pinning of the oracle's Davidson / LOBPCG restatements (davidson_cpu.cpp, lobpcg_cpu.cpp)
and their dense steps, on the CPU.

* random_col and eig_zhegv(_eval): against the reference's own compiled sources
  (oracle/_ref, built from util_random_cpu.cpp and eig_zhegv.cpp), bit for bit and to 1e-12.
* The solvers: LOBPCG's converged eigenvalues against a dense diagonalisation of the same
  Hamiltonian (1e-9 Ha); Davidson's Ritz values descending monotonically towards them and
  never below (min-max); and the first iteration of each against an independent
  Rayleigh-Ritz on the subspace the algorithm spans (span{X0, prec R0}), so a wrong block,
  sign or preconditioner placement fails.
"""
import numpy as np
import pytest

from conftest import oracle_globals


def _ref():
    from oracle import oracle as O
    if not O.Reference.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return O.Reference()


def test_random_col_matches_reference_bit_for_bit(oracle):
    R = _ref()
    for seed in (100, 101, 127, 7):
        assert np.array_equal(oracle.random_col(3001, seed), R.random_col(3001, seed))
    X = oracle.start_vectors(50, 4)
    assert np.array_equal(X[3], R.random_col(50, 103))


def test_synth_start_vectors_and_prec_equal_the_oracle(oracle):
    """The package's vectorised start block and preconditioner (used by bench.py) are the
    oracle's random_col / GenPrec restatements."""
    from pwh import synth
    assert np.array_equal(synth.start_vectors(700, 5), oracle.start_vectors(700, 5))
    x = np.linspace(0.0, 40.0, 101)
    for tau in (-10.0, 0.3):
        assert np.array_equal(synth.gen_prec(x, 0.4, 10.26, tau), oracle.gen_prec(x, 0.4, 10.26, tau))


def test_random_col_known_answer(oracle):
    """First element for seed 100: a1 = (100 * 214013 + 2531011) >> 16 & 32767 = 365,
    a2 = (365 * 214013 + 2531011) >> 16 & 32767 = 1230 -> (66/365, 31/1230)."""
    v = oracle.random_col(1, 100)[0]
    assert v == complex(66 / 365, 31 / 1230)


def _herm(n, rng):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return A + A.conj().T


def test_eig_zhegv_matches_reference(oracle):
    R = _ref()
    rng = np.random.default_rng(5)
    n = 14
    A = _herm(n, rng)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = M @ M.conj().T + n * np.eye(n)
    e, E, nn, small = oracle.eig_zhegv_eval(A, B)
    er, Er = R.eig_zhegv_eval(A, B)
    assert nn == n and small.size == 0
    assert np.abs(e - er).max() < 1e-12 * np.abs(er).max()
    # eigenvectors agree up to a phase per column (the spectrum is simple)
    for j in range(n):
        ph = np.vdot(E[:, j], Er[:, j])
        assert abs(abs(ph) / (np.linalg.norm(E[:, j]) * np.linalg.norm(Er[:, j])) - 1.0) < 1e-10
    Eo = R.eig_zhegv(A, B)
    assert np.abs(np.abs(np.sum(Eo.conj() * Er, axis=0)) - np.sum(np.abs(Er) ** 2, axis=0)).max() < 1e-9
    # generalized eigenproblem A E = B E diag(e), E^H B E = I
    assert np.abs(A @ E - B @ E * e).max() < 1e-10 * np.abs(e).max()
    assert np.abs(E.conj().T @ B @ E - np.eye(n)).max() < 1e-11


def test_eig_zhegv_reports_small_overlap_eigenvalues(oracle):
    """eig_zhegv.cpp:40-52: overlap eigenvalues <= 12.22e-16 leave their column unscaled and
    are reported; the others are normalised."""
    rng = np.random.default_rng(6)
    n = 8
    V = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    B = V @ V.conj().T                       # rank 5
    A = _herm(n, rng)
    e, E, nn, small = oracle.eig_zhegv_eval(A, B)
    assert nn == 5 and small.size == 3
    assert np.all(small <= oracle.ZHEGV_SMALL)


def test_gen_prec_limits(oracle):
    """GenPrec.cpp:11-24: the Teter-like form is 1 at G = 0 and ~ 1/(2x) for large
    x = |G|^2 (bohr^-2); the tau form is tau / (0.5 (XKG + 1) TPIBA2)."""
    p = oracle.gen_prec(np.array([0.0, 1e4]), 0.4, 2 * np.pi, -1.0)
    assert p[0] == 1.0
    assert abs(p[1] * 2e4 - 1.0) < 2e-4
    q = oracle.gen_prec(np.array([1.0, 3.0]), 0.5, 1.0, 0.2)
    assert np.allclose(q, [0.2 / (0.5 * 2.0 * 0.5), 0.2 / (0.5 * 4.0 * 0.5)], rtol=0, atol=1e-15)


@pytest.fixture(scope="module")
def tiny():
    """A 24^3 synthetic cell with 437 plane waves; the lowest 12 eigenvalues form two
    6-fold groups with a gap of 0.58 Ha above them."""
    from pwh import synth
    from oracle import oracle as O
    s = synth.make_system(10.26, 8.0, 1, (24, 24, 24), 12, name="tiny")
    g = oracle_globals(s)

    def H(X):
        return O.apply_ham_block(g, np.ascontiguousarray(X), s.vloc, s.ngw, nthreads=8)

    Hd = H(np.eye(s.ngw, dtype=np.complex128)).T
    w = np.linalg.eigvalsh(Hd)
    prec = O.gen_prec(s.gv["xkg"][:s.ngw], s.tpiba2, s.a, -1.0)
    return s, H, Hd, w, prec


def test_davidson_ritz_values_descend_to_the_dense_eigenvalues(tiny, oracle):
    """Each Davidson step is a Rayleigh-Ritz over a space that contains the current block,
    so the Ritz values never rise and never fall below the exact eigenvalues (min-max);
    they approach them (the method is a preconditioned block steepest descent and
    converges slowly on this deep synthetic potential)."""
    s, H, Hd, w, prec = tiny
    n = 12
    r = oracle.davidson(H, oracle.start_vectors(s.ngw, n), prec, max_iter=60)
    E = np.array(r.evals_history)
    assert np.all(np.diff(E, axis=0) <= 1e-11)
    assert np.all(E >= w[:n] - 1e-11)
    assert np.abs(E[-1, :6] - w[:6]).max() < 2e-3          # the lowest group converges first
    assert np.sum(E[-1] - w[:n]) < 0.05 * np.sum(E[0] - w[:n])


def test_lobpcg_converges_to_dense_eigenvalues(tiny, oracle):
    s, H, Hd, w, prec = tiny
    n = 12
    r = oracle.lobpcg(H, oracle.start_vectors(s.ngw, n), prec, max_iter=200)
    assert r.reason == 2 and r.iterations < 200
    assert np.abs(r.evals - w[:n]).max() < 1e-9
    assert np.all(np.array(r.evals_history) >= w[:n] - 1e-11)
    # the final block spans the lowest 12 eigenvectors
    _, V = np.linalg.eigh(Hd)
    P = V[:, :n] @ V[:, :n].conj().T
    Qx, _ = np.linalg.qr(r.X.T)
    assert np.abs(Qx @ Qx.conj().T - P).max() < 1e-5


def test_davidson_first_step_is_rayleigh_ritz_on_x_and_preconditioned_residual(tiny, oracle):
    s, H, Hd, w, prec = tiny
    n = 6
    X0 = oracle.start_vectors(s.ngw, n)
    r = oracle.davidson(H, X0, prec, max_iter=1)
    X = oracle.ortho_qr(X0).T
    lam = np.real(np.sum(X.conj() * (Hd @ X), axis=0))
    Rm = (X * lam - Hd @ X) * prec[:, None]
    Qb, _ = np.linalg.qr(np.concatenate([X, Rm], axis=1))
    ritz = np.linalg.eigvalsh(Qb.conj().T @ Hd @ Qb)
    assert np.abs(r.evals - ritz[:n]).max() < 1e-10


def test_lobpcg_first_step_is_rayleigh_ritz_on_x_and_preconditioned_residual(tiny, oracle):
    """After the first iteration X holds the lowest Ritz vectors of span{X0, prec W0}; the
    lambda of the second iteration are their Rayleigh quotients."""
    s, H, Hd, w, prec = tiny
    n = 6
    X0 = oracle.start_vectors(s.ngw, n)
    r = oracle.lobpcg(H, X0, prec, max_iter=2)
    X = oracle.ortho_qr(X0).T
    T1 = X.conj().T @ Hd @ X
    Wm = (Hd @ X - X @ T1) * prec[:, None]
    Qb, _ = np.linalg.qr(np.concatenate([X, Wm], axis=1))
    ritz = np.linalg.eigvalsh(Qb.conj().T @ Hd @ Qb)
    assert r.iterations == 2 and r.reason == 0
    assert np.abs(r.evals - ritz[:n]).max() < 1e-10
