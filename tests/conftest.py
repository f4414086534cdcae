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
    """The density G-vectors of a synthetic cell (GVectorsGen.cpp: |G|^2 <= 4 ecut for the
    wave-function cutoff ecut) with its structure factor and Gaussian local form factors
    (synthetic, as the cells are): returns an oracle GSpace."""
    from oracle import oracle as O
    from pwh import gvectors
    gv = gvectors(s.A, s.ecut, s.nr)
    ng, ngx = gv["ng"], gv["ngx"]
    in1, in2, in3 = (gv["in123"][k * ngx:k * ngx + ng].astype(np.int64) for k in range(3))
    nr = s.nr
    n1 = np.where(in1 < 0, in1 + 1 + nr[0], in1 + 1)
    n2 = np.where(in2 < 0, in2 + 1 + nr[1], in2 + 1)
    n3 = np.where(in3 < 0, in3 + 1 + nr[2], in3 + 1)
    B = np.linalg.inv(s.A).T * s.a                  # reciprocal vectors in 2pi/a
    Gv = np.stack([in1, in2, in3], axis=1) @ B
    g2 = np.sum(Gv * Gv, axis=1)
    tau = s.tau.reshape(-1, 3) / s.a
    sfac = np.sum(np.exp(-2j * np.pi * (Gv @ tau.T)), axis=1)[None, :]
    omega = abs(np.linalg.det(s.A))
    tpiba2 = (2 * np.pi / s.a) ** 2
    gg = g2 * tpiba2
    with np.errstate(divide="ignore"):
        vps = np.where(gg > 0, -4 * np.pi * 4.0 / (omega * np.where(gg > 0, gg, 1.0)) * np.exp(-0.25 * gg), 0.3)
    rhops = (-4.0 / omega * np.exp(-0.25 * gg))
    return O.GSpace(nr=tuple(nr), n1=n1, n2=n2, n3=n3, g2=g2, sfac=sfac, vps=vps[None, :],
                    rhops=rhops[None, :], tpiba2=tpiba2, omega=omega)
