/*
 * pwh.h -- C ABI of libpwh: MI355X (gfx950) Kohn-Sham H|psi> application and
 * Chebyshev-filtered subspace iteration, a drop-in for the hot path of
 * f-fathurrahman/ffr-pspw-dft-c.
 *
 * Conventions (identical to the reference's globals and arguments):
 *   - complex arrays are interleaved (re, im) doubles, i.e. C99 `double complex`;
 *   - wave-function blocks are column-major [ngw x nstates] with ld = ngw (the
 *     reference's vin/vout/X, ApplyHam_block.cpp:34-61);
 *   - N123 is 1-based, grid index i1 + NR1*(i2 + NR2*i3) (GkVectorsGen.cpp:30);
 *   - EIGR(ig, ia, is) at ig + ld*(ia + NAX*is); PKG_A(l, ig, is) at
 *     l + NLMAX*(ig + ld*is); WNL(is, l) at is + NSP*l (common_pspw.h:17-20).
 * All pointers are host pointers unless the name says "device".  Every call returns
 * PWH_OK (0) or a negative PWH_ERR_* code; pwh_last_error() describes the failure.
 * The library is not reentrant: one caller thread, one HIP stream (the reference's
 * own threading model, common_pspw.h globals).
 */
#ifndef PWH_H
#define PWH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PWH_OK 0
#define PWH_ERR_ARG -1          /* invalid argument (the reference aborts: nnr != NNR, lm_end > 4) */
#define PWH_ERR_HIP -2          /* HIP runtime failure (no device, launch failure, OOM) */
#define PWH_ERR_UNSUPPORTED -3  /* grid size without a compiled FFT plan */
#define PWH_ERR_LAPACK -4       /* factorisation failure (Cholesky / eigensolver info != 0) */

#define PWH_FP64 0    /* FP64 FFT, FP64 MFMA GEMMs */
#define PWH_FP32 1    /* FP32 FFT, FP32 MFMA GEMMs */
#define PWH_BF16X3 2  /* FP32 FFT, GEMMs on two-term (hi/lo) bf16 splits, FP32 accumulation */
#define PWH_BF16X6 3  /* FP32 FFT, GEMMs on three-term bf16 splits (six products, FP32-level
                         accuracy), FP32 accumulation */

typedef struct pwh_system pwh_system;

/* ---- library ------------------------------------------------------------ */
/* Select the HIP device and create the library stream (idempotent). */
int pwh_init(int device);
const char *pwh_last_error(void);
int pwh_device_count(int *count);
/* Self-test of the matrix-core lane layouts: kind 0 f64, 1 f32, 2 bf16. */
int pwh_selftest_mfma(int kind, double *max_err);
/* Whether the FFT passes can transform this grid: every length in 2..4096 (any factors; the
 * reference plans any size through FFTW, fft_fftw3.cpp:29,35). */
int pwh_grid_supported(int nr1, int nr2, int nr3);

/* ---- setup: producers of the reference's hot-path globals (host, FP64) ------ */
/* Fourier grid from the cutoff: Setup.cpp:20-34.  A = A1,A2,A3 (bohr), ecut in Ry. */
int pwh_fft_grid(const double A[9], double ecut, int nr[3]);
/* NGX, NGWX: GVectorsGen0.cpp:5-63 + Setup.cpp:63-66. xk: XK(1:3,1:nkpt) in 2pi/a. */
int pwh_gvec_sizes(const double A[9], double ecut, const int nr[3], const double *xk, int nkpt,
                   int *ngx, int *ngwx);
/* G and G+k maps: GVectorsGen.cpp:5-137 (with sort1.cpp:6-55) and GkVectorsGen.cpp:5-38.
 * in123 [3*ngx] (IN1,IN2,IN3), ngw [nkpt], igk/xkg/n123 [ngwx*nkpt], ggk [3*ngwx*nkpt]. */
int pwh_gvectors(const double A[9], double ecut, const int nr[3], const double *xk, int nkpt,
                 int ngx, int ngwx, int *ng, int *in123, int *ngw, int *igk, double *xkg,
                 unsigned int *n123, double *ggk);
/* EIGR = exp(-i G.tau): PhaseFactor.cpp:5-80.  tau = TAU(1:3,1:nax,1:nsp) bohr.
 * eigr [ngwx*nax*nsp*nkpt]. */
int pwh_phase_factor(const double A[9], const int nr[3], int nsp, const int *na, int nax,
                     const double *tau, int nkpt, const int *ngw, int ngwx, const int *igk,
                     int ngx, const int *in123, double *eigr /* complex */);

/* No counterpart in the reference (it stores EIGR and forms the projector rows per call,
 * ApplyHam_block.cpp:88-95): where the projectors of the FP32 H|psi> GEMMs come from.
 * mode 0 (default): the beta rows stored at system creation; mode 1: generated on the fly
 * inside the GEMMs from the factorised phases EI1 EI2 EI3 (PhaseFactor.cpp:41-77, built
 * from A and tau as pwh_phase_factor_gpu does: tau [3*nax*nsp] cartesian) and PKG_A.
 * FP64 and BF16 GEMMs, ortho and the grouped batch filter keep the stored rows. */
int pwh_projector_mode(pwh_system *sys, int mode, const double A[9], const double *tau);

/* ---- set-up on the GPU: phases, structure factor, form factors (FP64) ------------ */
/* PhaseFactor.cpp:5-80 computed on the GPU: same arguments and output as
 * pwh_phase_factor (agrees to rounding: the device sin/cos differ from libm by an ulp). */
int pwh_phase_factor_gpu(const double A[9], const int nr[3], int nsp, const int *na, int nax,
                         const double *tau, int nkpt, const int *ngw, int ngwx, const int *igk,
                         int ngx, const int *in123, double *eigr /* complex */);
/* StructureFactor.cpp:5-29 on the GPU: sfac [nsp*ng] complex (species fastest) for the
 * first ng density G-vectors of in123 [3*ngx] (IN1, IN2, IN3). */
int pwh_structure_factor(const double A[9], const int nr[3], int nsp, const int *na, int nax,
                         const double *tau, int ng, int ngx, const int *in123, double *sfac /* complex */);
/* The radial pseudopotential data ReadPseudopotentials.cpp:26-63 fills. */
typedef struct {
  int nsp;                    /* NSP */
  int maxr, nang;             /* leading dimension of r (MAXR) and channel count of psi/vion (MAXAM) */
  const int *mmax;            /* MMAX [nsp], odd */
  const double *clog;         /* CLOG [nsp]: log of the mesh ratio */
  const double *r;            /* R [maxr*nsp] */
  const double *psi, *vion;   /* PSI, VION [maxr*nsp*nang] */
  const double *zv, *rgauss;  /* ZV, RGAUSS [nsp] */
  const int *l_max, *l_loc;   /* L_MAX, L_LOC [nsp] */
} pwh_pseudo_desc;
/* NLFormFactor.cpp:5-133 on the GPU: wnl [nsp*nlmax] (WNL(is, m), species fastest),
 * pkg_a [nlmax*ngwx*nsp*nkpt] (PKG_A(m, ig, is, ik)), pkg [ngwx*nsp*nkpt*nlmax] (PKG(ig,
 * is, ik, m); may be null).  xkg [ngwx*nkpt], ggk [3*ngwx*nkpt], xk [3*nkpt].  A channel
 * above l = 3 -> PWH_ERR_UNSUPPORTED (the reference aborts). */
int pwh_nl_form_factor(const pwh_pseudo_desc *ps, double omega, double tpiba, int nkpt, const int *ngw,
                       int ngwx, const double *xkg, const double *ggk, const double *xk, int nlmax,
                       double *wnl, double *pkg, double *pkg_a);
/* FormFactor.cpp:5-72 on the GPU: vps, rhops [nsp*ng] (species fastest; RHOPS at G = 0 is
 * left as it was), eself.  g [ng] = |G|^2 in (2pi/a)^2.  L_LOC < 1 -> PWH_ERR_ARG. */
int pwh_form_factor(const pwh_pseudo_desc *ps, const int *na, double omega, double tpiba2, int ng,
                    const double *g, double *vps, double *rhops, double *eself);

/* ---- the device-resident system (one k-point) ------------------------------ */
typedef struct {
  int nr1, nr2, nr3;          /* NR1..NR3 */
  int ngw;                    /* NGW(ik) */
  int ngwx;                   /* NGWX: leading dimension of eigr/pkg_a */
  const unsigned int *n123;   /* N123(:,ik) [ngw], 1-based */
  const double *xkg;          /* XKG(:,ik) [ngw], |k+G|^2 in (2pi/a)^2 */
  double tpiba2;              /* TPIBA2 */
  int nsp, nax, nlmax;        /* NSP, NAX, NLMAX */
  const int *na;              /* NA [nsp] */
  const int *l_max;           /* L_MAX [nsp] */
  const int *l_loc;           /* L_LOC [nsp] */
  const double *wnl;          /* WNL [nsp*nlmax] */
  const double *eigr;         /* EIGR(:,:,:,ik) [ngwx*nax*nsp] complex */
  const double *pkg_a;        /* PKG_A(:,:,:,ik) [nlmax*ngwx*nsp] */
} pwh_system_desc;

/* Uploads the maps, phases and projectors; replaces the reference's per-call reads of
 * the globals N123, XKG, EIGR, PKG_A, WNL (ApplyHam_block.cpp:34-163). */
int pwh_system_create(const pwh_system_desc *desc, pwh_system **sys);
int pwh_system_destroy(pwh_system *sys);
int pwh_system_info(const pwh_system *sys, int *ngw, int *ncol, int *nxa, size_t *device_bytes);
/* No reference counterpart (FFTW picks its own plans): which implementation the system's FFT
 * passes use.  0 auto (the compiled per-length kernels when they cover the grid: 2,3,5-smooth
 * lengths 12..128 with NR1 = NR2; otherwise runtime plans), 1 compiled kernels (else
 * PWH_ERR_UNSUPPORTED), 2 runtime plans with one (i3, band) plane per workgroup, 3 runtime
 * plans with the plane pass split into i2 / i1 / i2 passes through HBM.  A plane that does
 * not fit one workgroup's LDS in a precision takes the split passes in any mode.  *used
 * (may be null) receives the mode in effect.  Results agree across modes to rounding. */
int pwh_fft_path(pwh_system *sys, int path, int *used);
/* The rho argument of ApplyHam_block (V_eff on the grid, EvalEnergy.cpp:128-131).
 * nnr != NR1*NR2*NR3 -> PWH_ERR_ARG (ApplyHam_block.cpp:18-21). */
int pwh_set_vloc(pwh_system *sys, const double *rho, int nnr);

/* ---- the hot path (host pointers, reference order) ------------------------- */
/* ApplyHam_block.cpp:5-172 (ApplyHam.cpp:5-142 for nstates = 1): vout = H vin. */
int pwh_apply_ham_block(pwh_system *sys, const double *vin, double *vout, int ngw_ik, int nstates,
                        int precision);
/* chebyfilt.cpp:6-74: degree-m scaled Chebyshev filter on X [ngw x nstates], in place. */
int pwh_chebyfilt(pwh_system *sys, double *X, int nstates, int degree, double lb, double ub,
                  int precision);
/* ortho_qr.cpp:3-34: orthonormalise the columns of X [nbasis x nstates] in place (FP64
 * CholeskyQR2 on the GPU; columns span the same space as Householder QR, column phases
 * follow the Cholesky convention R_jj > 0).  When X^H X is not numerically positive
 * definite (cond(X) above ~1e8, or X rank deficient) it runs Householder QR and the
 * explicit Q (rocSOLVER zgeqrf + zungqr), as the reference does. */
int pwh_ortho_qr(double *X, int nbasis, int nstates);
/* lanczos.cpp:8-101: k-step Lanczos (FP64) from v0 [ngw] (normalised in place);
 * T [ldT x k] column-major as the reference fills it, f [ngw] the final residual. */
int pwh_lanczos(pwh_system *sys, double *v0, int k, double *T, int ldT, double *f);
/* Rayleigh-Ritz block, ChebySCFLoop.cpp:289-311 (+ symmetrize lobpcg_cpu.cpp:15-27,
 * eig_zheevd.cpp:5-34): X orthonormal in, X <- X Q out, eig [nstates] ascending. */
int pwh_rayleigh_ritz(pwh_system *sys, double *X, int nstates, double *eig);
/* Residual report, ChebySCFLoop.cpp:415-433: res[j] = || H x_j - X (X^H H X)_sym e_j ||. */
int pwh_residual_norms(pwh_system *sys, const double *X, int nstates, double *res);

/* ---- density from the wave functions ----------------------------------------- */
/* EvalRhoPsi.cpp:83-157 on the GPU (FP64): for nstates states of eigvec [ld x nstates]
 * (reference order, ld >= ngw), rhoe[nnr] += sum_ist w[ist] / omega |psi_ist(r)|^2 with
 * psi(r) the unscaled inverse FFT of coefficients 1 .. ngw-1 (the reference's loop stops
 * one short, :117); w[ist] = WKPT(ik) * FOCC(ist, ik).  charge[ist] = sum |c|^2 over the
 * same coefficients (the reference aborts when |charge - 1| > 6e-3, :129-133; the caller
 * checks); *ekin = sum_ist w[ist] / 2 * sum XKG |c|^2 * TPIBA2, the reference's EKIN
 * increment (:154). */
int pwh_eval_rho_psi(pwh_system *sys, const double *eigvec, int ld, int nstates, const double *w,
                     double omega, double *rhoe, double *ekin, double *charge);

/* ---- SCF steps around the eigensolver (FP64, host arrays as the reference's) ---- */
/* The density G-vectors and local form factors EvalEnergy.cpp reads, and k-point 1's
 * sphere for the Kerker filter (KerkerRhoMix.cpp:26-35). */
typedef struct {
  int nr1, nr2, nr3;          /* NR1..NR3 */
  int ng;                     /* NG: density G-vectors, G = 0 first */
  const int *n1, *n2, *n3;    /* N1..N3 [ng]: 1-based grid indices (GVectorsGen.cpp:107-125) */
  const double *g2;           /* G [ng]: |G|^2 in (2pi/a)^2 */
  int nsp;                    /* NSP */
  const double *sfac;         /* SFAC [nsp*ng] complex, species fastest (StructureFactor.cpp) */
  const double *vps, *rhops;  /* VPS, RHOPS [nsp*ng], species fastest (FormFactor.cpp) */
  double tpiba2, omega;       /* TPIBA2, OMEGA */
  int ngw;                    /* NGW(1) */
  const unsigned int *n123;   /* N123(:,1) [ngw], 1-based */
  const double *xkg;          /* XKG(:,1) [ngw] */
} pwh_grid_desc;
typedef struct pwh_grid pwh_grid;
int pwh_grid_create(const pwh_grid_desc *desc, pwh_grid **grid);
int pwh_grid_destroy(pwh_grid *grid);
/* Energy terms of EvalEnergy.cpp:137-141: in ekin, enl, esr, eself; out the rest.
 * eh_electron / eh_pseudo: the electron-only and pseudo-charge-only Hartree sums the
 * reference forms (:96-100) times OMEGA.  n_negative: grid points below -0.1 (xc_PZ warns
 * for each). */
typedef struct {
  double ekin, enl, esr, eself;
  double etot, eht, epseu, exc, eh_electron, eh_pseudo;
  int n_negative;
} pwh_energy;
/* EvalEnergy.cpp:5-163: rhoe [nnr] the density in, V_eff = V_H + V_loc + V_xc out (the
 * forward transform of the density keeps its components outside the G set, :113-117). */
int pwh_eval_energy(pwh_grid *grid, double *rhoe, pwh_energy *e);
/* xc_PZ.cpp:7-63: Perdew-Zunger LDA; exc = sum rho (e_x + e_c) (not yet times OMEGA/NNR).
 * Points below -0.1 reuse the previous point's terms (n_negative counts them). */
int pwh_xc_pz(const double *rho, double *vxc, double *exc, int nnr, int *n_negative);
/* NormalizeRho.cpp:7-33: negatives to zero, scaled to nel electrons; diff = electron
 * count difference before scaling (the reference warns above 1e-5); sum < 1e-15 ->
 * PWH_ERR_ARG (the reference aborts). */
int pwh_normalize_rho(double *rho, int nnr, double omega, double nel, double *diff);
/* The density mixers with the reference's arguments and in-place updates:
 * SimpleRhoMix.cpp:5-20, KerkerRhoMix.cpp:5-55 and :57-82 (KerkerPrec), AndersonRhoMix.cpp:9-85
 * (rho_old, f [nnr x 3], g [nnr x 2]), PulayRhoMix.cpp:5-91 (rho_old, f [nnr x mixdim],
 * mixdim 2..7, beta 0.25 as the reference fixes it; the result is normalised to nel
 * electrons and nel_diff (may be null) gets NormalizeRho's count difference), BroydenRhoMix.cpp:10-107 (rho_old, f [nnr x 2], df, u [nnr x mixdim],
 * a [mixdim x mixdim], mixdim 2..8).  d: the RMS (Anderson: the squared norm) of the
 * density change, as each reference routine reports it. */
int pwh_mix_simple(double beta, double *rhoe, double *rho_old, int nnr, double *d);
int pwh_mix_kerker(pwh_grid *grid, double beta, double q0, double *rhoe, double *rho_old, double *d);
int pwh_kerker_prec(pwh_grid *grid, double *diff_rhoe, int idx1, double beta, double q0);
int pwh_mix_anderson(int iter, double beta, double *rhoe, double *rho_old, double *f, double *g, int nnr,
                     double *d);
int pwh_mix_pulay(pwh_grid *grid, int iter, double *rhoe, double *rho_old, double *f, int mixdim, double q0,
                  double nel, double *d, double *nel_diff);
int pwh_mix_broyden(int iter, double alpha, double w0, double *rho, double *rho_old, double *f, double *df,
                    double *u, double *a, int mixdim, int nnr, double *d);
/* No counterpart in the reference, whose SCF drivers keep the density, the potential and
 * the mixing history in host arrays (ChebySCFLoop.cpp:171-372): the same state held in HBM,
 * so an SCF step around pwh_resident_chfsi moves only scalars between host and device.
 * The steps are the host-array entry points' (same kernels, same order, same results):
 *   pwh_scf_state_density   EvalRhoPsi of the first nstates resident states -> the density
 *                           (EvalRhoPsi.cpp:83-157; w, omega, ekin, charge as in
 *                           pwh_eval_rho_psi; the density is overwritten, not added to)
 *   pwh_scf_state_mix       the density mixed in place with the held history (mixer 0:
 *                           SimpleRhoMix (beta), 1: KerkerRhoMix (beta, q0), 2: AndersonRhoMix
 *                           (beta), 4: BroydenRhoMix (alpha, w0; mixdim 2..8))
 *   pwh_scf_state_energy    EvalEnergy of the density: V_eff held, energy terms out
 *   pwh_scf_state_set_vloc  the held V_eff as the system's local potential
 * pwh_scf_state_set_rho sets the start density and resets the history as the reference's
 * arrays start (ChebySCFLoop.cpp:121-168: rho_old = the start density for the simple and
 * Kerker mixers; the Anderson and Broyden histories zero).  pwh_scf_state_get
 * copies the density and/or V_eff out (either may be null).  The grid and the systems
 * must share NNR. */
typedef struct pwh_scf_state pwh_scf_state;
int pwh_scf_state_create(pwh_grid *grid, int mixer, int mixdim, pwh_scf_state **state);
int pwh_scf_state_destroy(pwh_scf_state *state);
int pwh_scf_state_set_rho(pwh_scf_state *state, const double *rho);
int pwh_scf_state_get(pwh_scf_state *state, double *rho, double *veff);
int pwh_scf_state_density(pwh_system *sys, pwh_scf_state *state, int nstates, const double *w, double omega,
                          double *ekin, double *charge);
int pwh_scf_state_mix(pwh_scf_state *state, int iter, double beta, double q0, double alpha, double w0, double *d);
int pwh_scf_state_energy(pwh_scf_state *state, pwh_energy *e);
int pwh_scf_state_set_vloc(pwh_system *sys, pwh_scf_state *state);
/* EvalENL.cpp:5-49 for one k-point: fnl [nproj x nstates] complex (row = projector in
 * species, atom, lm order; may be null) = beta^H C0, enl = WKPT sum FOCC |fnl|^2 WNL.
 * c0 [ld x nstates] in reference order, ld >= ngw. */
int pwh_eval_enl(pwh_system *sys, const double *c0, int ld, int nstates, const double *focc, double wkpt,
                 double *fnl, double *enl);

/* ---- adjacent dense routines (no system needed) ------------------------------ */
/* fft_fftw3.cpp:13-46: in-place 3D DFT of data [nr3][nr2][nr1] complex (x fastest).
 * inverse != 0: exp(+i G.r), unscaled (FFTW_BACKWARD); inverse == 0: exp(-i G.r) then
 * 1/(nr1 nr2 nr3) (FFTW_FORWARD + zdscal).  FP64 on the GPU; each length must have a
 * length in 2..4096 (compiled kernels for 2,3,5-smooth 12..128, runtime plans otherwise),
 * else PWH_ERR_UNSUPPORTED. */
int pwh_fft3d(double *data, int nr1, int nr2, int nr3, int inverse);
/* eig_zheevd.cpp:5-34: Hermitian eigensolve of A [ldA x n] column-major, upper triangle
 * referenced ('V','U'); A <- eigenvectors, lambda [n] ascending.  FP64 on the GPU: the
 * library's own solver for n <= 128 (rocSOLVER zheevd when it flags a non-finite value),
 * rocSOLVER zheevd above; info != 0 -> PWH_ERR_LAPACK. */
int pwh_eig_zheevd(double *A, int ldA, double *lambda, int n);
/* No counterpart in the reference: pwh_eig_zheevd's solver (n <= 128) for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  A_device [ldA x n] (upper triangle read, eigenvectors out),
 * lambda_device [n] and info_device [1] are device pointers; the work is enqueued on
 * `stream` (a hipStream_t; null: the library stream, which needs pwh_init) and the call
 * returns without waiting.  Its scratch is a fixed 530 KB device array per GPU, shared by
 * every call on that GPU: calls that overlap on one GPU race.  n > 128 -> PWH_ERR_ARG.
 * *info_device = 0, or nonzero when a non-finite value appeared (pwh_eig_zheevd would then
 * rerun rocSOLVER; this entry allocates nothing and reports it instead). */
int pwh_eig_zheevd_noalloc(double *A_device, int ldA, double *lambda_device, int n, int *info_device,
                           void *stream);

/* ---- ChFSI outer loop (eigen part of ChebySCFLoop.cpp:271-311) -------------- */
typedef struct {
  int degree;            /* CHEBYPOL_DEGREE */
  int niter;             /* outer iterations */
  int precision;         /* filter precision (PWH_FP64 / PWH_FP32 / PWH_BF16X3 / PWH_BF16X6) */
  int lanczos_k;         /* steps of the FP64 Lanczos bound estimate */
  int bounds_mode;       /* 0: ub = max Ritz + ||f|| once, lb = largest RR value after each RR
                            1: the reference's rules (ChebySCFLoop.cpp:176-196,355-368) */
  double lb, ub;         /* in: used when lanczos_k == 0; out: final bounds */
  double t_filter_s, t_ortho_s, t_rr_s, t_lanczos_s, t_total_s;   /* out: wall times */
  long long band_applies;                                          /* out */
} pwh_chfsi_opts;
/* X [ngw x nstates] start vectors in, Ritz vectors out; eig/res [nstates]. v0 [ngw]: the
 * Lanczos start vector (ChebySCFLoop.cpp:102-107), normalised in place. */
int pwh_chfsi(pwh_system *sys, double *X, int nstates, double *v0, pwh_chfsi_opts *opts,
              double *eig, double *res);

/* ---- block eigensolvers of the SCF driver ------------------------------------ */
/* Events passed to pwh_diag_opts.report (the lines the reference prints, in order). */
#define PWH_DIAG_ITER 1           /* top of an outer iteration: a = iteration, b = nconv, value = mean residual */
#define PWH_DIAG_SMALL_OVERLAP 2  /* eig_zhegv.cpp:47-49: overlap eigenvalue a (1-based) of b is <= 12.22e-16 (value) */
#define PWH_DIAG_RANK 3           /* eig_zhegv.cpp:51-54: a linearly independent vectors of b */
#define PWH_DIAG_EIG_FAILED 4     /* zheevd failed with info a on an order-b matrix (the reference prints, goes on) */
typedef void (*pwh_diag_report_fn)(void *ctx, int event, int a, int b, double value);
typedef struct {
  int max_iter;                 /* in: outer iterations (the reference: max(30, nstates/2), or MAX_DIAG_ITER) */
  int nempty;                   /* in: NEMPTY; stop when nconv >= nstates - nempty */
  pwh_diag_report_fn report;    /* in: optional progress / warning callback */
  void *report_ctx;
  int iterations;               /* out: outer iterations run */
  int nconv;                    /* out: residual norms below 1e-6 at exit */
  int reason;                   /* out: 0 iterations exhausted, 1 nconv test, 2 mean-residual test */
  double t_total_s;             /* out: wall time */
  long long band_applies;       /* out: H|psi> band applications */
} pwh_diag_opts;
/* davidson_cpu.cpp:9-247 on the GPU (FP64): X [ngw x nstates] holds the start vectors
 * (the reference's random_col columns; orthonormalised here as by ortho_qr) and receives
 * the final block; prec [ngw] the preconditioner (GenPrec.cpp); evals [nstates] the
 * Ritz values, res [nstates] the residual norms of the final block.  Reduced problems of
 * order 2 nstates are solved as eig_zhegv_eval does. */
int pwh_davidson(pwh_system *sys, double *X, int nstates, const double *prec, pwh_diag_opts *opts,
                 double *evals, double *res);
/* lobpcg_cpu.cpp:32-389 on the GPU (FP64), same arguments: lambda [nstates] the
 * eigenvalues of X^H H X and res [nstates] the residual norms from the last convergence
 * test.  A Cholesky failure of a block overlap returns PWH_ERR_LAPACK (the reference
 * prints "ERROR calculating Cholesky decomposition" and aborts). */
int pwh_lobpcg(pwh_system *sys, double *X, int nstates, const double *prec, pwh_diag_opts *opts,
               double *lambda, double *res);

/* ---- device-resident entry points (inputs already in HBM; used for timing) --- */
/* Copy a block into / out of the system's resident FP64 wave-function buffer.  The
 * host-pointer entry points above stage their input in that same buffer, so a call to
 * any of them replaces the resident block. */
int pwh_resident_set(pwh_system *sys, const double *X, int nstates);
int pwh_resident_get(pwh_system *sys, double *X, int nstates);
/* One H|psi> over the resident block (nb bands) in the given precision. */
int pwh_resident_apply(pwh_system *sys, int nstates, int precision);
/* Launch shapes of H|psi> for this system and precision.  With *target_blocks == 0 (or a
 * null pointer) it times H|psi> of the resident block (nstates bands) and keeps the fastest
 * setting (the default unless another is faster by more than 1%) of: the projections'
 * split-K block target (256, 512, 1024, 2048; default 512); then, every combination timed,
 * the LDS stage depth of the projection and of the expansion GEMM (16, 32, and 64 for FP32;
 * default 16; FP64 and FP32 only) with the plane-pass workgroups per CU while the nonlocal
 * GEMMs overlap the FFT passes (no cap, 3, 2, 1; default no cap); the split-K target once
 * more at those; and, for grids off the compiled kernels, the radix order of the runtime FFT
 * plans (pwh_tuning_get reports all of them).  It returns the target in *target_blocks and
 * the time per H|psi> in *best_ms (both optional).  With *target_blocks in [64, 65536] it
 * sets that target without timing (nstates unused), e.g. one found by an earlier run.
 * Every later H|psi> of the system in that precision uses the settings.  No counterpart in
 * the reference: it tunes launch shapes of this library (the reference's GEMMs are cuBLAS
 * calls, ApplyHam.cpp). */
int pwh_autotune(pwh_system *sys, int nstates, int precision, int *target_blocks, double *best_ms);
/* No counterpart in the reference: the settings pwh_autotune chose (or the defaults) for a
 * precision: out[0] the projections' split-K block target, out[1] / out[2] the LDS stage
 * depth of the projection / expansion GEMMs, out[3] the radix-order variant of the runtime
 * FFT plans (-1: the grid runs on the compiled per-length kernels), out[4] the cap on
 * plane-pass workgroups per CU beside the overlapping GEMMs (0: none). */
int pwh_tuning_get(pwh_system *sys, int precision, int out[5]);
/* No counterpart in the reference: sets the five settings pwh_tuning_get reports, without
 * timing (e.g. those of an earlier pwh_autotune run, so a profiled run repeats a tuned one).
 * 0 selects a knob's default; out-of-range values return PWH_ERR_ARG and change nothing. */
int pwh_tuning_set(pwh_system *sys, int precision, const int in[5]);
/* Chebyshev filter of the resident block (FP32 modes convert in and out once). */
int pwh_resident_chebyfilt(pwh_system *sys, int nstates, int degree, double lb, double ub,
                           int precision);
/* pwh_chfsi (ChebySCFLoop.cpp:271-311) on the first nstates resident states: the Ritz
 * vectors stay resident (and its FP32 copy is refreshed); eig/res come back. */
int pwh_resident_chfsi(pwh_system *sys, int nstates, double *v0, pwh_chfsi_opts *opts, double *eig,
                       double *res);
/* pwh_eval_rho_psi (EvalRhoPsi.cpp:83-157) on the first nstates resident states (the
 * resident block is left as it is). */
int pwh_resident_eval_rho_psi(pwh_system *sys, int nstates, const double *w, double omega,
                              double *rhoe, double *ekin, double *charge);
/* pwh_eval_enl (EvalENL.cpp:5-49) on the first nstates resident states. */
int pwh_resident_eval_enl(pwh_system *sys, int nstates, const double *focc, double wkpt, double *fnl,
                          double *enl);
int pwh_synchronize(void);

/* ---- many uneven systems through one captured filter ------------------------ */
/* No counterpart in the reference (it runs one system per process): a batch of
 * independent systems (different grids, plane-wave and band counts) whose resident
 * blocks are Chebyshev-filtered together.  Every system's chain of passes runs on its
 * own stream; mode 1 captures all of them, with fork/join edges, into one hipGraph
 * that is replayed while degree, precision and bounds stay the same. */
typedef struct pwh_batch pwh_batch;
int pwh_batch_create(pwh_system *const *systems, const int *nstates, int nsys, pwh_batch **batch);
int pwh_batch_destroy(pwh_batch *batch);
/* Upload / download system isys's block [ngw x nstates] (reference order). */
int pwh_batch_set(pwh_batch *batch, int isys, const double *X);
int pwh_batch_get(pwh_batch *batch, int isys, double *X);
/* chebyfilt.cpp:6-74 on every resident block; lb/ub [nsys].  mode 0: systems one after
 * the other on the library stream; mode 1: one hipGraph with a branch per system; mode 2:
 * grouped, every pass of every H|psi> of the filter is one launch over all systems that
 * run the same kernel (the systems of one grid share the compiled per-length kernel;
 * systems on runtime plans share one launch), with one descriptor per system for its
 * maps, plans, buffers and scalars; the launches are captured into one hipGraph.  Mode 2
 * takes PWH_FP64 and PWH_FP32 and systems with at least two states whose plane pass does
 * not need the split path, and gives each system its one-at-a-time result bit for bit;
 * anything else returns PWH_ERR_UNSUPPORTED. */
int pwh_batch_chebyfilt(pwh_batch *batch, int degree, const double *lb, const double *ub,
                        int precision, int mode);

/* ---- per-kernel timing with HIP events on the library stream ---------------- */
/* While timing is on, H|psi> runs every kernel on the library stream (no overlap of the
 * nonlocal GEMMs with the FFT passes), so each event pair times one kernel alone. */
int pwh_timing_enable(int on);
int pwh_timing_reset(void);
/* name: "zinv", "plane", "zfwd", "proj", "expand", "hpsi" (H|psi> passes), or "rho_sums",
 * "rho_zinv", "rho_plane", "rho" (the device part of pwh_eval_rho_psi: permutation and
 * per-state sums, i3 transform, plane pass with the band sum, all three), or "energy" (the
 * two device sections of pwh_eval_energy: transforms and G-space sums, then xc_PZ).
 * total_ms / count of launches. */
int pwh_timing_get(const char *name, double *total_ms, int *count);

#ifdef __cplusplus
}
#endif
#endif /* PWH_H */
