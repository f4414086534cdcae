"""Synthetic Si-diamond systems for the benchmark and parity configurations.

Each system is what the reference's Setup/PrepareSCF would leave in its globals for a
simple-cubic cell of Si atoms on a diamond lattice at the Gamma point, with two
synthetic ingredients in place of pseudopotential files:

  * local potential  V(r) = sum_atoms sum_images -2 exp(-|r - R|^2 / 2)  Ha
    on the grid r = (i1/NR1, i2/NR2, i3/NR3) * a (the rho argument of ApplyHam_block);
  * 4 projectors per atom (s and p, i.e. L_MAX = 3, L_LOC = 3 -> lm_end = 4),
    PKG_A(l, G) = exp(-(0.8 |G|)^2 / 2) * {1, Gx/|G|, Gy/|G|, Gz/|G|}, |G| in 1/bohr,
    p components zero at G = 0 (the unnormalised real-harmonic convention of
    NLFormFactor.cpp:56-86), WNL = (+1, -0.5, -0.5, -0.5) Ha.

The G maps and phases come from libpwh's host setup (pwh_gvectors / pwh_phase_factor,
bit-exact with GVectorsGen / GkVectorsGen / PhaseFactor).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Globals, gvectors, phase_factor

# 8-atom cubic diamond cell, fractional coordinates in the order of works/Si8/INPUT:25-32
DIAMOND8 = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.75, .25, .75], [.25, .75, .75], [.75, .75, .25]])

WNL_SP = (1.0, -0.5, -0.5, -0.5)
PROJ_WIDTH = 0.8     # bohr
V_DEPTH = -2.0       # Ha
V_WIDTH = 1.0        # bohr

# name -> (a [bohr], ecut [Ry], cells per edge, FFT grid, bands)
CONFIGS = {
    "si8": (10.26, 15.0, 1, (32, 32, 32), 16),
    "si64": (20.52, 20.0, 2, (64, 64, 64), 128),
    "si216": (30.8, 20.0, 3, (96, 96, 96), 432),
}


def diamond_fractional(ncell):
    """Fractional positions of an ncell^3 supercell of the 8-atom diamond cell."""
    out = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                out.append((DIAMOND8 + np.array([i, j, k])) / ncell)
    return np.concatenate(out, axis=0)


def gaussian_vloc(a, nr, tau):
    """V on the grid, x fastest (flat [nr1*nr2*nr3]); orthorhombic cell of edges a (a
    number: cubic), images +-1 along each axis."""
    edges = np.broadcast_to(np.asarray(a, np.float64), (3,))
    axes = [np.arange(n) / n * edges[d] for d, n in enumerate(nr)]
    V = np.zeros((nr[2], nr[1], nr[0]))
    for R in tau:
        g = []
        for d in range(3):
            x = axes[d][:, None] - R[d] + np.array([-1.0, 0.0, 1.0])[None, :] * edges[d]
            g.append(np.exp(-x * x / (2.0 * V_WIDTH ** 2)).sum(axis=1))
        V += g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return (V_DEPTH * V).reshape(-1)


@dataclass
class SynthSystem:
    name: str
    a: float
    ecut: float
    nr: tuple
    nbands: int
    A: np.ndarray
    tau: np.ndarray       # (natoms, 3) bohr
    gv: dict
    eigr: np.ndarray
    pkg_a: np.ndarray
    wnl: np.ndarray
    vloc: np.ndarray
    tpiba2: float

    @property
    def ngw(self):
        return int(self.gv["ngw"][0])

    @property
    def natoms(self):
        return int(self.tau.shape[0])

    def globals(self) -> Globals:
        ngw = self.ngw
        return Globals(nr=self.nr, ngw=ngw, ngwx=self.gv["ngwx"], n123=self.gv["n123"][:ngw],
                       xkg=self.gv["xkg"][:ngw], tpiba2=self.tpiba2,
                       na=np.array([self.natoms], np.int32), nax=self.natoms,
                       l_max=np.array([3], np.int32), l_loc=np.array([3], np.int32), nlmax=4,
                       wnl=self.wnl, eigr=self.eigr, pkg_a=self.pkg_a)

    def psi(self, nbands=None, seed=0):
        """complex128 i.i.d. N(0,1) coefficients from the seed, QR-orthonormalised (nb, ngw)."""
        nb = self.nbands if nbands is None else nbands
        rng = np.random.default_rng(seed)
        X = rng.standard_normal((nb, self.ngw)) + 1j * rng.standard_normal((nb, self.ngw))
        Q, _ = np.linalg.qr(X.T, mode="reduced")
        return np.ascontiguousarray(Q.T)


def make_system(a, ecut, ncell, nr, nbands, name="custom", natoms=None, frac=None, edges=None):
    """A cubic cell of edge a, or the orthorhombic cell of the given edges (then ALAT =
    edges[0], Setup.cpp:39, and a is ignored)."""
    if edges is not None:
        edges = np.asarray(edges, np.float64)
        a = float(edges[0])
    else:
        edges = np.array([a, a, a], np.float64)
    A = np.diag(edges).astype(np.float64)
    if frac is None:
        frac = diamond_fractional(ncell)
        if natoms is not None:
            frac = frac[:natoms]
    tau = np.asarray(frac, np.float64) * edges[None, :]
    gv = gvectors(A, ecut, nr)
    nat = tau.shape[0]
    eigr = phase_factor(A, nr, [nat], nat, tau.reshape(1, nat, 3), gv)
    ngw, ngwx = int(gv["ngw"][0]), gv["ngwx"]
    tpiba = 2.0 * np.pi / a
    xkg = gv["xkg"][:ngw]
    t = np.sqrt(xkg)
    q = np.exp(-((PROJ_WIDTH * t * tpiba) ** 2) / 2.0)
    ggk = gv["ggk"][:3 * ngw].reshape(ngw, 3)
    pkg = np.zeros((ngwx, 4))
    pkg[:ngw, 0] = q
    nz = t >= 1.0e-4
    for d in range(3):
        pkg[:ngw, 1 + d][nz] = q[nz] * ggk[nz, d] / t[nz]
    pkg_a = np.ascontiguousarray(pkg.reshape(-1))     # l fastest: [l + 4*ig]
    wnl = np.array(WNL_SP, np.float64)                 # WNL(1,l) at 0 + 1*l
    vloc = gaussian_vloc(edges, nr, tau)
    return SynthSystem(name=name, a=a, ecut=ecut, nr=tuple(nr), nbands=nbands, A=A, tau=tau, gv=gv,
                       eigr=eigr, pkg_a=pkg_a, wnl=wnl, vloc=vloc, tpiba2=tpiba * tpiba)


def config(name):
    a, ecut, ncell, nr, nb = CONFIGS[name]
    return make_system(a, ecut, ncell, nr, nb, name=name)


# Cells on the grids the reference's own runs use, which the compiled FFT lengths do not
# cover: name -> (edges [bohr], ecut [Ry], grid, atoms, bands).  The grids are those
# Setup.cpp:20-34 picks for the edges and cutoff (28^3: works/Si8/log_15Ry_lobpcg:54;
# 22 x 12 x 16: works/NH3/log_davidson_cpu_10Ry:53; 40 x 40 x 82), or given as a user grid
# (70 x 70 x 140: works/TiO2_anatase001/INPUT:13-14).  Atoms sit at seeded positions.
GRID_CELLS = {
    "si8_28": ((10.266, 10.266, 10.266), 15.0, (28, 28, 28), 8, 16),
    "ortho_22x12x16": ((10.0, 5.0, 7.2), 10.0, (22, 12, 16), 3, 8),
    "slab_40x40x82": ((19.0, 19.0, 40.0), 10.0, (40, 40, 82), 6, 16),
    "anatase_70x70x140": ((19.1464, 19.1464, 40.0), 30.0, (70, 70, 140), 12, 8),
}


def grid_cell(name, seed=11):
    edges, ecut, nr, nat, nb = GRID_CELLS[name]
    if name == "si8_28":
        frac = DIAMOND8
    else:
        frac = np.random.default_rng(seed).random((nat, 3))
    return make_system(None, ecut, 1, nr, nb, name=name, frac=frac, edges=edges)


def smooth_grid(n):
    """Smallest even 2,3,5-smooth size >= n (the FFT sizes the library compiles)."""
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    n = max(n, 2)
    while n % 2 or not smooth(n):
        n += 1
    return n


def uneven_batch(count=24, seed=2026, setup_grid=False):
    """Independent cubic cells with a ~ U[8,18] bohr and ecut ~ U[15,20] Ry.

    The grid is the Setup.cpp:20-34 size rounded up to an even 2,3,5-smooth length
    (setup_grid=True: the Setup.cpp size itself, any even length, as the reference runs).
    The cell holds round(8 (a/10.26)^3) atoms (Si density) at seeded random fractional
    positions, with 4 projectors per atom and 2*natoms + 4 bands (clipped to 8..96).
    """
    from . import fft_grid
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        a = float(rng.uniform(8.0, 18.0))
        ecut = float(rng.uniform(15.0, 20.0))
        n = fft_grid(np.diag([a, a, a]), ecut)[0]
        if not setup_grid:
            n = smooth_grid(n)
        nat = max(1, int(round(8.0 * (a / 10.26) ** 3)))
        frac = rng.random((nat, 3))
        nb = int(min(96, max(8, 2 * nat + 4)))
        out.append(make_system(a, ecut, 1, (n, n, n), nb, name=f"cell{i}", frac=frac))
    return out


def start_vectors(nbasis, nstates):
    """The start block of the reference's Davidson and LOBPCG (davidson_cpu.cpp:47-51,
    lobpcg_cpu.cpp:80-84): column ist from random_col with seed 100 + ist - 1
    (util_random_cpu.cpp:5-25: each part is (a % 100 + 1) / a after one step of
    a <- ((214013 a + 2531011) >> 16) & 32767).  Rows = states; the columns run side by
    side."""
    a = np.arange(100, 100 + nstates, dtype=np.int64)
    out = np.empty((nstates, 2 * nbasis))
    for i in range(2 * nbasis):
        a = ((a * 214013 + 2531011) >> 16) & 32767
        out[:, i] = (a % 100 + 1) / a
    return np.ascontiguousarray(out.view(np.complex128))


def gen_prec(xkg, tpiba2, alat, tau):
    """GenPrec.cpp:5-27, the preconditioner Diagon_*_cpu hands to the block solvers:
    y / (y + 16 x^4), y = 27 + x (18 + x (12 + 8 x)), x = |G|^2 for tau <= 0, else
    tau / (0.5 (XKG + 1) TPIBA2)."""
    xkg = np.asarray(xkg, np.float64)
    if tau <= 0.0:
        x = xkg * (2.0 * np.pi / alat) ** 2
        y = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
        return y / (y + 16.0 * x ** 4)
    return tau / (0.5 * (xkg + 1.0) * tpiba2)


def density_gspace(s):
    """The density G-vectors of a synthetic cell (GVectorsGen.cpp: |G|^2 <= 4 ecut for the
    wave-function cutoff ecut), their structure factor for the cell's atoms and Gaussian
    local form factors (synthetic, like the cells' projectors).  Returns the arrays
    pwh.Grid takes: n1, n2, n3 (1-based grid indices), g2 (|G|^2 in (2pi/a)^2), sfac
    [1][ng], vps [1][ng], rhops [1][ng], tpiba2, omega."""
    from . import gvectors
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
    safe = np.where(gg > 0, gg, 1.0)
    vps = np.where(gg > 0, -4 * np.pi * 4.0 / (omega * safe) * np.exp(-0.25 * gg), 0.3)
    rhops = -4.0 / omega * np.exp(-0.25 * gg)
    return dict(n1=n1, n2=n2, n3=n3, g2=g2, sfac=sfac, vps=vps[None, :], rhops=rhops[None, :],
                tpiba2=tpiba2, omega=omega)
