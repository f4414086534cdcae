"""CPU checks of the boundary: libpwh.so loads and exports every entry point declared in
include/pwh.h, the shim exports the reference's mangled C++ symbols, and the library's
host setup (G maps, phases) equals the reference bit for bit (no GPU compute here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "pwh.h")

# Itanium-mangled names of the reference's hot-path routines (common_pspw.h:176-178,187,222)
# and the adjacent ones the shim also replaces (lanczos :185-186, eig_zheevd :217 and its
# eig_zheevd.cpp:37 overload, fft_fftw3 :348, EvalRhoPsi :179-180), as g++ -std=gnu++98
# mangles the reference's declarations
REF_SYMBOLS = ["_Z14ApplyHam_blockPCdS0_iiPdii", "_Z8ApplyHamPCdS0_iiPdi",
               "_Z9chebyfiltiPCdiddPdi", "_Z8ortho_qrPCdii", "_Z7lanczosiPCdiPdiS0_S1_i",
               "_Z10eig_zheevdPCdiPdi", "_Z10eig_zheevdPCdii", "_Z9fft_fftw3PCdiiib",
               "_Z10EvalRhoPsiiPdPCdi", "_Z10EvalRhoPsiiPdPCdii", "_Z12davidson_cpuiPdPCdiiS_iS_",
               "_Z10lobpcg_cpuiPdPCdiiS_iS_", "_Z10EvalEnergyPd", "_Z5xc_PZPdS_Rdi", "_Z7EvalENLv",
               "_Z12NormalizeRhoPdidd", "_Z12SimpleRhoMixdPdS_iRd", "_Z12KerkerRhoMixddPdS_iRd",
               "_Z10KerkerPrecPdiidd", "_Z14AndersonRhoMixidPdS_S_S_iRd", "_Z11PulayRhoMixiPdS_S_iiRd",
               "_Z13BroydenRhoMixiddPdS_S_S_S_S_iiRd", "_Z12NLFormFactorv", "_Z10FormFactorv"]


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pwh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import pwh
    L = pwh.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert sorted(pwh.EXPORTS) == names


def test_shim_exports_reference_signatures():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "libpwh_shim.so")],
                         capture_output=True, text=True, check=True).stdout
    for sym in REF_SYMBOLS:
        assert re.search(r"\sT\s" + re.escape(sym) + r"$", out, flags=re.M), sym


def test_shim_loads_with_reference_globals():
    ctypes.CDLL(os.path.join(PKG, "libpwh_refglobals.so"), mode=ctypes.RTLD_GLOBAL)
    shim = ctypes.CDLL(os.path.join(PKG, "libpwh_shim.so"))
    for sym in REF_SYMBOLS:
        assert hasattr(shim, sym)


@pytest.mark.parametrize("name", ["si8_8ry", "si8_15ry", "si8_mixed"])
def test_host_setup_equals_reference_fixtures(name):
    import pwh
    z = np.load(os.path.join(ROOT, "tests", "golden", f"ref_maps_{name}.npz"))
    A, nr = z["A"], tuple(int(x) for x in z["nr"])
    gv = pwh.gvectors(A, float(z["ecut"]), nr)
    n = int(z["ngw"])
    assert gv["ngw"][0] == n
    assert np.array_equal(gv["n123"][:n], z["n123"])
    assert np.array_equal(gv["igk"][:n], z["igk"])
    assert np.array_equal(gv["xkg"][:n], z["xkg"])
    tau = z["tau"]
    nat = tau.shape[0]
    e = pwh.phase_factor(A, nr, [nat], nat, tau[None], gv).reshape(nat, gv["ngwx"])[:, :n]
    assert np.array_equal(e, z["eigr"])


@pytest.mark.parametrize("a,ecut,ncell,nr", [(10.266, 8.0, 1, (20, 20, 20)), (20.52, 20.0, 2, (64, 64, 64)),
                                             (30.8, 20.0, 3, (96, 96, 96)), (10.0, 12.0, 1, (24, 30, 36)),
                                             (9.0, 16.0, 1, (28, 28, 28))])
def test_host_setup_equals_oracle(oracle, a, ecut, ncell, nr):
    import pwh
    from pwh import synth
    A = np.diag([a, a, a])
    tau = synth.diamond_fractional(ncell) * a
    nat = len(tau)
    g1 = pwh.gvectors(A, ecut, nr)
    g2 = oracle.gvectors(A, ecut, nr)
    for k in ("in123", "igk", "xkg", "n123", "ggk", "ngw"):
        assert np.array_equal(g1[k], g2[k]), k
    assert (g1["ng"], g1["ngx"], g1["ngwx"]) == (g2["ng"], g2["ngx"], g2["ngwx"])
    e1 = pwh.phase_factor(A, nr, [nat], nat, tau[None], g1)
    e2 = oracle.phase_factor(A, nr, [nat], nat, tau[None], g2)
    assert np.array_equal(e1, e2)
    assert pwh.fft_grid(A, ecut) == oracle.fft_grid(A, ecut)


def test_config_sizes():
    """Exact plane-wave counts of the benchmark cells."""
    import pwh
    for a, ecut, nr, ngw in [(10.26, 15.0, (32,) * 3, 1021), (20.52, 20.0, (64,) * 3, 13133),
                             (30.8, 20.0, (96,) * 3, 43867)]:
        assert pwh.gvectors(np.diag([a] * 3), ecut, nr)["ngw"][0] == ngw


def test_grid_support_table():
    import pwh
    for n in (24, 30, 32, 36, 40, 48, 50, 54, 60, 64, 72, 80, 90, 96):
        assert pwh.grid_supported((n, n, n))
    # any length 2..4096 (runtime plans): prime factors, non-square planes, long axes
    for nr in ((22, 22, 22), (32, 30, 32), (22, 12, 16), (70, 70, 140), (41, 43, 47), (2, 2, 4096)):
        assert pwh.grid_supported(nr), nr
    for nr in ((1, 8, 8), (8, 8, 4097), (0, 8, 8)):
        assert not pwh.grid_supported(nr), nr


def test_no_device_is_a_clean_error():
    import pwh
    if pwh.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(pwh.PwhError) as e:
        pwh.init(0)
    assert e.value.code == pwh.PWH_ERR_HIP


def test_shim_aborts_on_nnr_mismatch_like_the_reference():
    """ApplyHam_block.cpp:18-21: the shim prints the same message and aborts before any
    device work (run in a child process)."""
    import subprocess
    import sys
    code = (
        "import ctypes as C, os, numpy as np\n"
        f"G = C.CDLL(os.path.join({PKG!r}, 'libpwh_refglobals.so'), mode=C.RTLD_GLOBAL)\n"
        f"S = C.CDLL(os.path.join({PKG!r}, 'libpwh_shim.so'))\n"
        "C.c_int.in_dll(G, 'NNR').value = 1000\n"
        "f = getattr(S, '_Z14ApplyHam_blockPCdS0_iiPdii')\n"
        "f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]\n"
        "f(None, None, 1, 1, None, 999, 1)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "nnr != NNR : 999 1000" in r.stdout


def test_shim_eval_rho_psi_aborts_on_wrong_charge_like_the_reference():
    """EvalRhoPsi.cpp:117-133: a state whose charge over coefficients 1 .. NGW-1 is off by
    more than 6e-3 stops the program with the reference's two lines; the shim checks the
    charges on the host before any device work (run in a child process)."""
    import subprocess
    import sys
    code = (
        "import ctypes as C, os, numpy as np\n"
        f"G = C.CDLL(os.path.join({PKG!r}, 'libpwh_refglobals.so'), mode=C.RTLD_GLOBAL)\n"
        f"S = C.CDLL(os.path.join({PKG!r}, 'libpwh_shim.so'))\n"
        "for n, v in (('NX', 2), ('NGWX', 4), ('NKPT', 1)): C.c_int.in_dll(G, n).value = v\n"
        "ngw = np.array([4], np.int32); c0 = np.zeros(8, np.complex128)\n"
        "C.c_void_p.in_dll(G, 'NGW').value = ngw.ctypes.data\n"
        "C.c_void_p.in_dll(G, 'C0').value = c0.ctypes.data\n"
        # state 1: charge 1 over the first three coefficients; state 2: 0.99 (+ a last
        # coefficient that the check leaves out)
        "ev = np.array([0.6, 0.8, 0, 5, 0.99 ** 0.5, 0, 0, 5], np.complex128)\n"
        "f = getattr(S, '_Z10EvalRhoPsiiPdPCdi')\n"
        "f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]\n"
        "f(1, None, ev.ctypes.data, 4)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "Wrong charge density in G-space:       0.9900000000" in r.stdout
    assert "for k-point 1 and states 2" in r.stdout
