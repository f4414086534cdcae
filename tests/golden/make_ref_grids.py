"""Writes ref_example_grids.json: the FFT grid of every example input under the reference's
works/ directory, as the reference picks it (Setup.cpp:20-34: n = int(2 sqrt(ECUT |a_i|^2) / pi)
+ 2, rounded up to even, unless the input gives fft_grid), with the lattice and cutoff it
came from, plus the grids the reference's own run logs print for two inputs that have
since changed.  Data only (numbers read from the INPUT files); run from the repository root:

    python tests/golden/make_ref_grids.py [/root/reference]
"""
import json
import math
import os
import sys


def parse_input(path):
    """latvec (3 rows), ecut and the optional fft_grid of a reference INPUT file."""
    lines = [l.strip() for l in open(path)]
    out = {}
    for i, l in enumerate(lines):
        if l == "latvec":
            out["A"] = [float(x) for r in lines[i + 1:i + 4] for x in r.split()[:3]]
        elif l == "ecut":
            out["ecut"] = float(lines[i + 1].split()[0])
        elif l == "fft_grid":
            out["fft_grid"] = [int(x) for x in lines[i + 1].split()[:3]]
    return out


def setup_grid(A, ecut):
    nr = []
    for d in range(3):
        a = A[3 * d:3 * d + 3]
        n = int(2.0 * math.sqrt(ecut * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2])) / math.pi) + 2
        nr.append(n + 1 if n % 2 else n)
    return nr


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    works = os.path.join(ref, "works")
    cases = {}
    for name in sorted(os.listdir(works)):
        f = os.path.join(works, name, "INPUT")
        if not os.path.isfile(f):
            continue
        d = parse_input(f)
        if "A" not in d or "ecut" not in d:
            continue
        rule = setup_grid(d["A"], d["ecut"])
        cases[name] = {"A": d["A"], "ecut": d["ecut"], "setup_grid": rule,
                       "grid": d.get("fft_grid", rule), "source": f"works/{name}/INPUT"}
    logs = {"Si8_15Ry": ("works/Si8/log_15Ry_lobpcg", [28, 28, 28]),
            "NH3_10Ry": ("works/NH3/log_davidson_cpu_10Ry", [22, 12, 16])}
    for k, (src, grid) in logs.items():
        txt = open(os.path.join(ref, src)).read()
        assert f"Fourier grid: ({grid[0]:8d},{grid[1]:8d},{grid[2]:8d})" in txt, (src, grid)
        cases[k] = {"grid": grid, "source": src}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_example_grids.json")
    with open(out, "w") as fh:
        json.dump(cases, fh, indent=1, sort_keys=True)
    print(f"{len(cases)} grids -> {out}")


if __name__ == "__main__":
    main()
