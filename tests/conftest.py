import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ffr-pspw-dft-c_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpwh on the GPU)")


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


def oracle_globals(s):
    """The oracle's view of a synthetic system (same arrays the GPU path gets)."""
    from oracle import oracle as O
    g = s.globals()
    return O.Globals(nr=s.nr, ngwx=g.ngwx, n123=s.gv["n123"], xkg=s.gv["xkg"], tpiba2=g.tpiba2,
                     na=g.na, nax=g.nax, l_max=g.l_max, l_loc=g.l_loc, nlmax=g.nlmax, wnl=g.wnl,
                     eigr=g.eigr, pkg_a=g.pkg_a)


@pytest.fixture(scope="session")
def si8():
    from pwh import synth
    return synth.config("si8")


@pytest.fixture(scope="session")
def gpu_si8(si8):
    import pwh
    pwh.init(0)
    sysm = pwh.System(si8.globals())
    sysm.set_vloc(si8.vloc)
    yield sysm
    sysm.close()


def density_gspace(s):
    """The synthetic cell's density G-vectors and form factors (pwh.synth.density_gspace)
    as an oracle GSpace."""
    from oracle import oracle as O
    from pwh import synth
    d = synth.density_gspace(s)
    return O.GSpace(nr=tuple(s.nr), **d)
