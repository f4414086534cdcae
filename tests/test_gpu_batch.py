"""Many uneven systems through one captured filter (pwh_batch_*).

Every system's result must equal the same filter run on that system alone (bitwise:
same kernels, same inputs), for the sequential schedule and for each of several
replays of the captured hipGraph with eager work in between; one system is also
checked against the oracle (1e-11)."""
import numpy as np
import pytest

from conftest import oracle_globals, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cells():
    from pwh import synth
    return synth.uneven_batch(count=6, seed=7)


def test_uneven_batch_is_uneven(cells):
    grids = {c.nr[0] for c in cells}
    assert len(grids) >= 3
    assert len({c.ngw for c in cells}) >= len(cells) - 1
    assert len({c.nbands for c in cells}) == len(cells)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_filter_matches_single_system(cells, oracle, prec):
    import pwh
    pwh.init(0)
    systems = []
    for c in cells:
        s = pwh.System(c.globals())
        s.set_vloc(c.vloc)
        systems.append(s)
    X0 = [c.psi(c.nbands, seed=i) for i, c in enumerate(cells)]
    lb = np.array([-1.0 + 0.1 * i for i in range(len(cells))])
    ub = np.array([40.0 + i for i in range(len(cells))])
    single = [s.chebyfilt(x, 5, l, u, prec) for s, x, l, u in zip(systems, X0, lb, ub)]
    B = pwh.Batch(systems, [c.nbands for c in cells])
    for mode in (0, 1, 1, 1):
        for i, x in enumerate(X0):
            B.set(i, x)
        B.chebyfilt(5, lb, ub, prec, mode)
        pwh.synchronize()
        for i in range(len(cells)):
            assert np.array_equal(B.get(i), single[i]), (mode, i)
        # eager work on the same systems between replays
        systems[0].apply_ham_block(X0[0][:2])
    # a retuned projection target changes buffers and grids: the graph is captured again,
    # so its replay equals the sequential schedule at the new target
    for s in systems:
        s.autotune(0, prec, target=2048)
    seq = []
    for mode in (0, 1, 1):
        for i, x in enumerate(X0):
            B.set(i, x)
        B.chebyfilt(5, lb, ub, prec, mode)
        pwh.synchronize()
        out = [B.get(i) for i in range(len(cells))]
        if mode == 0:
            seq = out
        for i in range(len(cells)):
            assert np.array_equal(out[i], seq[i]), ("retuned", mode, i)
            assert relerr(out[i], single[i]) < (1e-13 if prec == "fp64" else 1e-5)
    if prec == "fp64":
        c = cells[2]
        ref = oracle.chebyfilt(oracle_globals(c), X0[2], 5, lb[2], ub[2], c.vloc, c.ngw)
        assert relerr(single[2], ref) < 1e-11
    B.close()
    for s in systems:
        s.close()
