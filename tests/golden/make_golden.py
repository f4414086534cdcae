"""Writes the golden fixtures of tests/golden/ from the reference's own compiled code.

Run in the container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden.py

ref_maps_<name>.npz : N123, IGK, XKG (first NGW entries) and EIGR (nat x NGW) from the
                      reference's GVectorsGen0 / GVectorsGen / sort1 / GkVectorsGen /
                      PhaseFactor, for a few cells.
ref_lapack.npz      : the reference's ortho_qr (zgeqrf + zungqr) and eig_zheevd outputs
                      (MKL in this container) on seeded inputs, and on the 4x4 Hermitian
                      matrix of sandbox/t_zheevd.cpp:41-69.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ffr-pspw-dft-c_amd"))

from oracle.oracle import Reference  # noqa: E402

DIAMOND8 = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.75, .25, .75], [.25, .75, .75], [.75, .75, .25]])

CELLS = {
    # works/Si8/INPUT:19-35 (8 Ry, grid 20)
    "si8_8ry": (10.266, 8.0, (20, 20, 20)),
    # the Si8 parity configuration (15 Ry, grid 32)
    "si8_15ry": (10.26, 15.0, (32, 32, 32)),
    # a non-cubic grid with odd and mixed-radix extents
    "si8_mixed": (10.0, 12.0, (24, 30, 36)),
}


def main():
    R = Reference()
    for name, (a, ecut, nr) in CELLS.items():
        A = np.diag([a, a, a])
        tau = DIAMOND8 * a
        r = R.setup(A, ecut, nr, [8], 8, tau[None])
        n = int(r["ngw"][0])
        np.savez(os.path.join(HERE, f"ref_maps_{name}.npz"), A=A, ecut=ecut, nr=np.array(nr),
                 tau=tau, ngw=n, n123=r["n123"][:n], igk=r["igk"][:n], xkg=r["xkg"][:n],
                 eigr=r["eigr"].reshape(8, r["ngwx"])[:, :n])
        print(name, "NGW", n)
    rng = np.random.default_rng(2012)
    X = rng.standard_normal((6, 40)) + 1j * rng.standard_normal((6, 40))
    Q = R.ortho_qr(X)
    T = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    T = T + T.conj().T
    w, _ = R.eig_zheevd(T)
    vals = [(1, 0), (2, -1), (3, -1), (4, -1), (2, 1), (2, 0), (3, -2), (4, -2),
            (3, 1), (3, 2), (3, 0), (4, -3), (4, 1), (4, 2), (4, 3), (4, 0)]
    colmajor = np.array([complex(*v) for v in vals])
    A4 = colmajor.reshape(4, 4).T                 # A4[i, j] = A[i + 4 j]
    w4, _ = R.eig_zheevd(A4)
    np.savez(os.path.join(HERE, "ref_lapack.npz"), X=X, Q=Q, T=T, w=w, A4=A4, w4=w4)
    print("lapack", w4)


if __name__ == "__main__":
    main()
