"""A Kohn-Sham SCF loop shaped like ChebySCFLoop.cpp:171-372, composed from the library's
entry points: it runs the whole cycle (ChFSI step with the current V_eff, density,
nonlocal energy, mixing, EvalEnergy) on the GPU, so the mixed-precision filter can be
checked against the FP64 one on the converged total energy.

Per iteration, as the reference does:

    set V_eff -> ChFSI (filter, CholeskyQR, Rayleigh-Ritz; ChebySCFLoop.cpp:271-311)
    -> EvalRhoPsi on the lowest nocc states (:313) -> EvalENL (:314)
    -> mix (Simple / Kerker / Anderson / Broyden, :316-339) -> EvalEnergy (:342)
    -> |dE| test (:350)

What the caller supplies instead of the reference's set-up:

* the start density (the reference superposes atomic charges, FormFactorAtomic.cpp);
* ESR and ESELF (EwaldEnergy.cpp is not part of this library), constant offsets of EHT and
  so of ETOT (EvalEnergy.cpp:142-144);
* the spectral bounds follow the library's ChFSI rule (first call: ub = largest Lanczos
  Ritz value + ||f||; after each Rayleigh-Ritz step lb = its largest value).  The
  reference's update (ChebySCFLoop.cpp:355-368) adds the squared induced norm of a
  10-step Lanczos matrix (ChebySCFLoop.cpp:26-32), which puts ub tens of times above the
  spectrum and leaves a degree-10 filter almost flat on the occupied states.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import FP64, PRECISIONS, mix_anderson, mix_broyden, mix_simple


@dataclass
class ScfResult:
    etot: float
    energies: dict
    eig: np.ndarray              # Rayleigh-Ritz values of the last step
    X: np.ndarray                # Ritz vectors of the last step (nstates, ngw)
    rho: np.ndarray              # mixed density of the last step
    iterations: int
    converged: bool
    history: list = field(default_factory=list)   # (iteration, ETOT, |dE|, drho, lb, ub, precision)
    t_chfsi_s: float = 0.0
    t_total_s: float = 0.0


def cheby_scf(sysm, grid, X0, nocc, focc, omega, rho0, degree=10, precision=FP64, mixer="anderson",
              beta=0.3, max_iter=60, conv=1e-10, lanczos_k=20, v0=None, esr=0.0, eself=0.0,
              mixdim=4, kerker_q0=1.0, chfsi_iters=1, seed=1, fp64_refine=None, resident=True,
              autotune=False, device_state=True):
    """Runs the SCF loop from start vectors X0 (nstates, ngw) and density rho0 [nnr].
    nocc: occupied states (the first nocc Ritz vectors, ascending); focc [nocc]: their
    occupations times the k-point weight.  mixer: simple, kerker, anderson or broyden
    (IMIX 0/1/2/4 of the reference).  chfsi_iters: ChFSI outer iterations per SCF step (the
    reference runs one).  v0: the Lanczos start vector; by default uniform [0, 1) real and
    imaginary parts from `seed`, like the reference's rand() vector (ChebySCFLoop.cpp:102-107)
    -- a structured vector such as all ones stays in one symmetry subspace of the cell, its
    Krylov space runs out within a few steps and the bound estimate turns to noise.
    fp64_refine: with a mixed-precision filter, the |dE| at which the loop switches the
    filter to FP64 for the remaining iterations.  An FP32-filtered subspace carries
    FP32-sized noise, so the converged energy wanders by ~1e-7 relative to its size between
    iterations; the last iterations in FP64 remove it (None: no switch).  resident: keep the
    block in HBM for the whole loop (pwh_resident_chfsi / _eval_rho_psi / _eval_enl);
    otherwise every step passes the host block in and out.  autotune: before the first
    ChFSI step, pick the projection GEMMs' split-K targets for this block (pwh_autotune) for
    FP64 (Rayleigh-Ritz, the filter's last step) and the filter precision; needs
    resident=True.  device_state (with resident=True): hold the density, V_eff and the
    mixing history in HBM as well (pwh_scf_state_*), so a step moves only scalars between
    host and device; the steps and their results are those of the host-array entry points.
    Stops when
    |ETOT - ETOT_previous| <= conv (SCF_CONV_CRIT) or after max_iter iterations."""
    import time
    t0 = time.perf_counter()
    nnr = grid.nnr
    rho = np.array(rho0, np.float64).reshape(-1)
    if rho.size != nnr:
        raise ValueError("rho0 does not match the grid")
    focc = np.ascontiguousarray(focc, np.float64)
    if focc.size != nocc:
        raise ValueError("focc needs nocc entries")
    use_state = resident and device_state
    if mixer not in ("anderson", "broyden", "simple", "kerker"):
        raise ValueError("mixer: simple, kerker, anderson or broyden")
    if use_state:
        from . import ScfState
        st = ScfState(grid, mixer, mixdim)
        st.set_rho(rho)
        state = None
    elif mixer == "anderson":
        state = [np.zeros((3, nnr)), np.zeros((3, nnr)), np.zeros((2, nnr))]
    elif mixer == "broyden":
        # every history array starts at zero, as the reference allocates them
        # (ChebySCFLoop.cpp:137-149); only simple and Kerker mixing start from rho_old = rho0
        state = [np.zeros((2, nnr)), np.zeros((2, nnr)), np.zeros((mixdim, nnr)), np.zeros((mixdim, nnr)),
                 np.zeros((mixdim, mixdim))]
    else:
        state = [rho.copy()]
    if use_state:
        e = st.energy(0.0, 0.0, esr, eself)
    else:
        veff, e = grid.eval_energy(rho, 0.0, 0.0, esr, eself)
    etot_old = e["etot"]
    X = np.array(X0, np.complex128)
    if v0 is None:
        rng = np.random.default_rng(seed)
        v0 = rng.random(X.shape[1]) + 1j * rng.random(X.shape[1])
    lb = ub = 0.0
    hist, t_chfsi, converged, it = [], 0.0, False, 0
    eig = None
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    nst = X.shape[0]
    if autotune and not resident:
        raise ValueError("autotune times the resident block: use resident=True")
    if resident:
        sysm.resident_set(X)
    for it in range(1, max_iter + 1):
        if use_state:
            st.set_vloc(sysm)
        else:
            sysm.set_vloc(veff)
        if autotune and it == 1:
            for p in sorted({FP64, prec}):
                sysm.autotune(nst, p)
        ta = time.perf_counter()
        ck = dict(degree=degree, niter=chfsi_iters, precision=prec, lanczos_k=lanczos_k if it == 1 else 0,
                  bounds_mode=0, v0=v0, lb=lb, ub=ub)
        if resident:
            eig, _, stc = sysm.resident_chfsi(nst, **ck)
        else:
            eig, X, _, stc = sysm.chfsi(X, **ck)
        t_chfsi += time.perf_counter() - ta
        lb, ub = stc["lb"], stc["ub"]
        if use_state:
            ekin, _ = st.density(sysm, nocc, focc, omega)
            enl, _ = sysm.resident_eval_enl(nocc, focc, 1.0)
            d = st.mix(it, beta=beta, q0=kerker_q0, alpha=0.25, w0=0.01)
            e = st.energy(ekin, enl, esr, eself)
        else:
            if resident:
                rnew, ekin, _ = sysm.resident_eval_rho_psi(nocc, focc, omega)
                enl, _ = sysm.resident_eval_enl(nocc, focc, 1.0)
            else:
                rnew, ekin, _ = sysm.eval_rho_psi(X[:nocc], focc, omega)
                enl, _ = sysm.eval_enl(X[:nocc], focc, 1.0)
            rho = rnew
            if mixer == "anderson":
                d = mix_anderson(it, beta, rho, *state)
            elif mixer == "broyden":
                d = mix_broyden(it, 0.25, 0.01, rho, *state, mixdim)
            elif mixer == "kerker":
                d = grid.mix_kerker(beta, kerker_q0, rho, state[0])
            else:
                d = mix_simple(beta, rho, state[0])
            veff, e = grid.eval_energy(rho, ekin, enl, esr, eself)
        detot = abs(e["etot"] - etot_old)
        etot_old = e["etot"]
        hist.append((it, e["etot"], detot, d, lb, ub, prec))
        if detot <= conv:
            converged = True
            break
        if fp64_refine is not None and prec != FP64 and detot <= fp64_refine:
            prec = FP64
    if resident:
        X = sysm.resident_get(nst)
    if use_state:
        rho, _ = st.get()
        st.close()
    return ScfResult(etot=e["etot"], energies=e, eig=eig, X=X, rho=rho, iterations=it, converged=converged,
                     history=hist, t_chfsi_s=t_chfsi, t_total_s=time.perf_counter() - t0)
