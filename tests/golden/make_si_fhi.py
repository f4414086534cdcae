"""Write tests/golden/si_fhi.npz: the radial silicon pseudopotential data of the
reference's example input works/Si8/Si.fhi (FHI format as ReadPseudopotentials.cpp:26-63
reads it: ZV and the channel count, ten lines of Gaussian parameters, then per channel
MMAX, the mesh ratio, and MMAX lines of index, r, u(r), V(r)).  Data only; run in a tree
that has the reference checked out."""
import os
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/works/Si8/Si.fhi"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si_fhi.npz")


def main():
    lines = open(SRC).read().splitlines()
    zv, nang = lines[0].split()[:2]
    zv, nang = float(zv), int(nang)
    k = 11
    mmax, ratio, r, psi, vion = [], [], [], [], []
    for _ in range(nang):
        m, q = lines[k].split()[:2]
        m = int(m)
        mmax.append(m)
        ratio.append(float(q))
        rows = np.array([[float(x) for x in ln.split()[1:4]] for ln in lines[k + 1:k + 1 + m]])
        r.append(rows[:, 0])
        psi.append(rows[:, 1])
        vion.append(rows[:, 2])
        k += 1 + m
    np.savez_compressed(OUT, zv=zv, nang=nang, mmax=np.array(mmax, np.int32), ratio=np.array(ratio),
                        r=np.array(r), psi=np.array(psi), vion=np.array(vion))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
