"""GPU parity of H|psi> (ApplyHam_block / ApplyHam) against the FP64 oracle.

Tolerances: FP64 path 1e-12 relative (per band, 2-norm) as the reference-parity bar;
FP32 path 2e-6 and BF16x3 path 2e-5 relative (single-precision FFT / GEMM rounding);
BF16x6 (three-term bf16 splits) at the FP32 bound, 2e-6.
"""
import numpy as np
import pytest

from conftest import oracle_globals, relerr

pytestmark = pytest.mark.gpu


def band_relerr(a, b):
    return max(relerr(a[i], b[i]) for i in range(a.shape[0]))


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_mfma_lane_layout(kind):
    import pwh
    pwh.init(0)
    assert pwh.selftest_mfma(kind) == 0.0


def test_apply_ham_block_fp64_matches_oracle(si8, gpu_si8, oracle):
    psi = si8.psi(16, seed=1)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    out = gpu_si8.apply_ham_block(psi, "fp64")
    assert band_relerr(out, ref) < 1e-12


def test_apply_ham_single_band(si8, gpu_si8, oracle):
    psi = si8.psi(1, seed=2)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    out = gpu_si8.apply_ham_block(psi[0], "fp64")
    assert relerr(out, ref) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 2e-5), ("bf16x6", 2e-6)])
def test_apply_ham_block_mixed_error(si8, gpu_si8, oracle, prec, tol):
    psi = si8.psi(16, seed=3)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    out = gpu_si8.apply_ham_block(psi, prec)
    assert band_relerr(out, ref) < tol


def test_hermitian_and_linear(si8, gpu_si8):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((2, si8.ngw)) + 1j * rng.standard_normal((2, si8.ngw))
    Ha = gpu_si8.apply_ham_block(a)
    lhs = np.vdot(a[0], Ha[1])
    rhs = np.conj(np.vdot(a[1], Ha[0]))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    c = 0.3 - 1.7j
    Hsum = gpu_si8.apply_ham_block(a[0] + c * a[1])
    assert relerr(Hsum[0], Ha[0] + c * Ha[1]) < 1e-13


def test_odd_band_count_and_ragged(si8, gpu_si8, oracle):
    # nb not a multiple of the 64-wide GEMM tile, and a block of 67 bands
    psi = si8.psi(67, seed=5)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw, nthreads=8)
    out = gpu_si8.apply_ham_block(psi)
    assert band_relerr(out, ref) < 1e-12


def test_local_only_system(si8, oracle):
    """NSP = 1 with lm_end = 0 (L_MAX = L_LOC = ... -> no nonlocal part): kinetic + local only."""
    import pwh
    g = si8.globals()
    g.l_max = np.array([1], np.int32)
    g.l_loc = np.array([1], np.int32)   # lm_end = 1 - 2 + 1 = 0
    s = pwh.System(g)
    s.set_vloc(si8.vloc)
    og = oracle_globals(si8)
    og.l_max, og.l_loc = g.l_max, g.l_loc
    psi = si8.psi(4, seed=6)
    ref = oracle.apply_ham_block(og, psi, si8.vloc, si8.ngw)
    assert band_relerr(s.apply_ham_block(psi), ref) < 1e-12
    s.close()


def test_error_paths(si8, gpu_si8):
    import pwh
    with pytest.raises(pwh.PwhError) as e:
        gpu_si8.set_vloc(si8.vloc[:-1])          # nnr != NNR (ApplyHam_block.cpp:18-21)
    assert e.value.code == pwh.PWH_ERR_ARG
    with pytest.raises(pwh.PwhError) as e:
        gpu_si8.apply_ham_block(np.zeros((1, si8.ngw - 1), np.complex128))
    assert e.value.code == pwh.PWH_ERR_ARG
    g = si8.globals()
    g.l_max = np.array([4], np.int32)            # lm_end = 16 - 6 + 1 > 4 (ApplyHam_block.cpp:69-73)
    with pytest.raises(pwh.PwhError) as e:
        pwh.System(g)
    assert e.value.code == pwh.PWH_ERR_ARG
    g = si8.globals()
    g.nr = (32, 32, 5000)                        # a length above 4096: no FFT plan
    with pytest.raises(pwh.PwhError) as e:
        pwh.System(g)
    assert e.value.code == pwh.PWH_ERR_UNSUPPORTED


def test_si64_fp64_subset_matches_oracle(oracle):
    """Full C1-sized system (64^3, 13133 PW, 256 projectors), 128 bands on the GPU; the
    oracle checks 6 of them, and H is checked Hermitian on the full block."""
    import pwh
    from pwh import synth
    s = synth.config("si64")
    pwh.init(0)
    g = pwh.System(s.globals())
    g.set_vloc(s.vloc)
    psi = s.psi(128, seed=7)
    out = g.apply_ham_block(psi)
    idx = [0, 1, 63, 64, 100, 127]
    ref = oracle.apply_ham_block(oracle_globals(s), psi[idx], s.vloc, s.ngw, nthreads=8)
    assert band_relerr(out[idx], ref) < 1e-12
    T = psi.conj() @ out.T          # psi^H H psi, 128 x 128
    assert np.abs(T - T.conj().T).max() < 1e-12 * np.abs(T).max()
    for prec, tol in (("fp32", 2e-6), ("bf16x3", 2e-5), ("bf16x6", 2e-6)):
        o2 = g.apply_ham_block(psi, prec)
        assert band_relerr(o2[idx], ref) < tol, prec
    g.close()


def test_autotune_targets_keep_results(oracle):
    """pwh_autotune: every split-K block target gives the default target's H|psi> up to the
    order of the split-K sums (FP64 1e-13, FP32 2e-6 relative per band), the timed search
    picks one of its candidates, and bad arguments are refused."""
    import pwh
    from pwh import synth
    s = synth.config("si64")
    pwh.init(0)
    g = pwh.System(s.globals())
    g.set_vloc(s.vloc)
    psi = s.psi(128, seed=9)
    base = {p: g.apply_ham_block(psi, p) for p in ("fp64", "fp32")}
    for t in (64, 256, 1024, 2048, 4096):
        for p, tol in (("fp64", 1e-13), ("fp32", 2e-6)):
            g.autotune(0, p, target=t)
            assert band_relerr(g.apply_ham_block(psi, p), base[p]) < tol, (t, p)
    g.resident_set(psi)
    for p in ("fp64", "fp32", "bf16x3"):
        t, ms = g.autotune(128, p)
        assert t in (256, 512, 1024, 2048) and ms > 0.0
    idx = [0, 77, 127]
    ref = oracle.apply_ham_block(oracle_globals(s), psi[idx], s.vloc, s.ngw, nthreads=8)
    assert band_relerr(g.apply_ham_block(psi)[idx], ref) < 1e-12
    for bad in (dict(nstates=129), dict(nstates=128, precision=4), dict(nstates=0, target=32)):
        with pytest.raises(pwh.PwhError) as e:
            g.autotune(**bad)
        assert e.value.code == pwh.PWH_ERR_ARG
    g.close()


def test_plane_cap_and_tuning_set(oracle):
    """Capping the plane-pass workgroups per CU (pwh_tuning_set, the fifth knob of
    pwh_autotune) only pads their LDS: at the Si64 shape, where the nonlocal GEMMs overlap the
    FFT passes, H|psi> is bitwise equal for every cap in FP64 and FP32, and still matches the
    oracle; pwh_tuning_get reads back what was set; out-of-range settings are refused and
    change nothing."""
    import pwh
    from pwh import synth
    s = synth.config("si64")
    pwh.init(0)
    g = pwh.System(s.globals())
    g.set_vloc(s.vloc)
    psi = s.psi(128, seed=13)
    base = {p: g.apply_ham_block(psi, p) for p in ("fp64", "fp32")}
    for cap in (3, 2, 1):
        for p in ("fp64", "fp32"):
            g.set_tuning(p, {"plane_blocks_per_cu": cap})
            assert g.tuning(p)["plane_blocks_per_cu"] == cap
            assert np.array_equal(g.apply_ham_block(psi, p), base[p]), (cap, p)
    idx = [0, 127]
    ref = oracle.apply_ham_block(oracle_globals(s), psi[idx], s.vloc, s.ngw, nthreads=8)
    assert band_relerr(base["fp64"][idx], ref) < 1e-12
    g.set_tuning("fp32", {"target_blocks": 1024, "proj_stage_depth": 32, "expand_stage_depth": 64,
                          "plane_blocks_per_cu": 3})
    t = g.tuning("fp32")
    assert (t["target_blocks"], t["proj_stage_depth"], t["expand_stage_depth"], t["plane_blocks_per_cu"]) == \
        (1024, 32, 64, 3)
    assert band_relerr(g.apply_ham_block(psi, "fp32")[idx], ref) < 2e-6
    for p, bad in (("fp32", {"plane_blocks_per_cu": 4}), ("fp64", {"expand_stage_depth": 64}),
                   ("fp32", {"target_blocks": 32}), ("bf16x3", {"proj_stage_depth": 16})):
        before = g.tuning(p)
        with pytest.raises(pwh.PwhError) as e:
            g.set_tuning(p, bad)
        assert e.value.code == pwh.PWH_ERR_ARG
        assert g.tuning(p) == before
    g.close()


def test_projectors_on_the_fly_match_stored(si8, oracle):
    """pwh_projector_mode 1: the FP32 GEMMs generate beta from the factorised phases and
    PKG_A; H|psi> agrees with the stored-projector result to FP32 rounding and with the
    oracle to the FP32 bound; FP64 keeps the stored rows (bitwise unchanged); mode 0
    restores the stored path bit for bit."""
    import pwh
    pwh.init(0)
    g = pwh.System(si8.globals())
    g.set_vloc(si8.vloc)
    psi = si8.psi(16, seed=21)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    st32 = g.apply_ham_block(psi, "fp32")
    st64 = g.apply_ham_block(psi, "fp64")
    g.projector_mode(1, si8.A, si8.tau)
    fly32 = g.apply_ham_block(psi, "fp32")
    assert band_relerr(fly32, st32) < 1e-6
    assert band_relerr(fly32, ref) < 2e-6
    assert np.array_equal(g.apply_ham_block(psi, "fp64"), st64)
    g.projector_mode(0)
    assert np.array_equal(g.apply_ham_block(psi, "fp32"), st32)
    with pytest.raises(pwh.PwhError) as e:
        g.projector_mode(2, si8.A, si8.tau)
    assert e.value.code == pwh.PWH_ERR_ARG
    g.close()


def test_autotune_full_search_keeps_results(si8, oracle):
    """pwh_autotune searches the split-K target, both GEMMs' LDS stage depths jointly with
    the plane-pass cap, the target again and (on the runtime FFT path) the plans' radix
    order; whatever it keeps, H|psi> still matches the
    oracle (FP64 1e-12, FP32 2e-6), and pwh_tuning_get reports settings from the
    candidate sets."""
    import pwh
    pwh.init(0)
    g = pwh.System(si8.globals())
    g.set_vloc(si8.vloc)
    psi = si8.psi(16, seed=22)
    ref = oracle.apply_ham_block(oracle_globals(si8), psi, si8.vloc, si8.ngw)
    g.resident_set(psi)
    assert g.tuning("fp64") == {"target_blocks": 512, "proj_stage_depth": 16, "expand_stage_depth": 16,
                                "fft_plan_variant": -1, "plane_blocks_per_cu": 0}
    assert g.fft_path("runtime") == "runtime"
    for p, tol, depths in (("fp64", 1e-12, (16, 32)), ("fp32", 2e-6, (16, 32, 64))):
        t, ms = g.autotune(16, p)
        tu = g.tuning(p)
        assert tu["target_blocks"] == t and t in (256, 512, 1024, 2048) and ms > 0
        assert tu["proj_stage_depth"] in depths and tu["expand_stage_depth"] in depths
        assert tu["fft_plan_variant"] in (0, 1, 2)
        assert tu["plane_blocks_per_cu"] in (0, 1, 2, 3)
        assert band_relerr(g.apply_ham_block(psi, p), ref) < tol, (p, tu)
    assert g.fft_path("auto") == "compiled"
    assert g.tuning("fp64")["fft_plan_variant"] == -1
    assert band_relerr(g.apply_ham_block(psi, "fp64"), ref) < 1e-12
    g.close()
