"""The CPU schedules the benchmark times beside the GPU path compute the same H|psi> as
the oracle (so their rates are rates of the same work)."""
import os
import sys

import numpy as np

from conftest import oracle_globals, relerr

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_scipy_schedule_matches_oracle(si8, oracle):
    import bench
    psi = si8.psi(16, seed=11)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    out = bench.ScipyHpsi(si8, 2).apply(psi, chunk=5)        # ragged last chunk
    assert max(relerr(out[i], ref[i]) for i in range(psi.shape[0])) < 1e-12


def test_oracle_threads_match_one_thread(si8, oracle):
    psi = si8.psi(7, seed=12)
    g = oracle_globals(si8)
    a = oracle.apply_ham_block(g, psi, si8.vloc, si8.ngw)
    b = oracle.apply_ham_block(g, psi, si8.vloc, si8.ngw, nthreads=3)
    assert np.array_equal(a, b)


def test_cpu_info_reports_threads():
    import bench
    info = bench.cpu_info()
    assert info["usable_threads"] >= 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with status 2 and says so, before
    touching the GPU (this runs without one)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
