// ref_harness.cpp -- drives the reference's own compiled setup routines
// (GVectorsGen0, GVectorsGen, sort1, GkVectorsGen, PhaseFactor), its LAPACK
// wrappers (ortho_qr, eig_zheevd, eig_zhegv, eig_zhegv_eval) and the start-vector
// generator of the Davidson / LOBPCG solvers (random_col), xc_PZ, the density mixers
// that need no FFT (simple, Anderson, Broyden), NormalizeRho and EvalENL, so the oracle
// can be checked against them.
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref/, never shipped).
//
// Setup() itself (Setup.cpp:5-145) also reads pseudopotential files and input
// files, so this harness sets the reference's globals directly and repeats the
// scalar lines of Setup.cpp:39-66 before calling the reference routines in the
// order Setup.cpp:128,139 and PrepareSCF.cpp:12 call them.
#include "common_pspw.h"

void random_col(double complex *A, int nelem, long a);   // util_random_cpu.cpp:15

extern "C" {

static int g_ready = 0;

static void free_all()
{
  free(NGW); NGW = NULL;
  free(G); G = NULL; free(GG); GG = NULL;
  free(N1); N1 = NULL; free(N2); N2 = NULL; free(N3); N3 = NULL;
  free(IN1); IN1 = NULL; free(IN2); IN2 = NULL; free(IN3); IN3 = NULL;
  free(IGK); IGK = NULL; free(XKG); XKG = NULL; free(N123); N123 = NULL; free(GGK); GGK = NULL;
  free(EI1); EI1 = NULL; free(EI2); EI2 = NULL; free(EI3); EI3 = NULL; free(EIGR); EIGR = NULL;
  free(NA); NA = NULL; free(TAU); TAU = NULL; free(XK); XK = NULL; free(WKPT); WKPT = NULL;
  g_ready = 0;
}

// A: A1,A2,A3 (bohr), ecut (Ry), nr: FFT grid, tau: TAU(1:3,1:nax,1:nsp) bohr,
// xk: XK(1:3,1:nkpt) in 2pi/a.
int ref_setup(const double *A, double ecut, const int *nr, int nsp, const int *na, int nax,
              const double *tau, int nkpt, const double *xk)
{
  if (g_ready) free_all();
  for (int i = 0; i < 3; i++) { A1[i] = A[i]; A2[i] = A[3+i]; A3[i] = A[6+i]; }
  ECUT = ecut;
  NR1 = nr[0]; NR2 = nr[1]; NR3 = nr[2]; NNR = NR1*NR2*NR3; USER_FFT_GRID = true;
  NSP = nsp; NAX = nax;
  NA = (int*)malloc(nsp*sizeof(int));
  for (int i = 0; i < nsp; i++) NA[i] = na[i];
  TAU = (double*)malloc(3*nax*nsp*sizeof(double));
  for (int i = 0; i < 3*nax*nsp; i++) TAU[i] = tau[i];
  NKPT = nkpt;
  XK = (double*)malloc(3*nkpt*sizeof(double));
  WKPT = (double*)malloc(nkpt*sizeof(double));
  for (int i = 0; i < 3*nkpt; i++) XK[i] = xk[i];
  for (int i = 0; i < nkpt; i++) WKPT[i] = 1.0/nkpt;

  // Setup.cpp:39-66
  ALAT = A1[0];
  OMEGA = determinant3x3(A1, A2, A3);
  CalcRecipLatt();
  NGW = (int*)malloc(NKPT*sizeof(int));
  TPIBA = 2.*M_PI/ALAT;
  TPIBA2 = TPIBA*TPIBA;
  GCUT = 4.*ECUT/TPIBA/TPIBA;
  GCUTW = GCUT/4.0;
  GVectorsGen0();
  NGWX = NGX/8 + 8;
  NGX = NGWX*8 + 8;
  NGWX = NGWX + 1;

  // AllocateArrays.cpp:13-38 (the arrays these routines touch)
  G = (double*)calloc(NGX, sizeof(double));
  GG = (double*)calloc((NGX+1)*3, sizeof(double));
  N1 = (int*)calloc(NGX, sizeof(int)); N2 = (int*)calloc(NGX, sizeof(int)); N3 = (int*)calloc(NGX, sizeof(int));
  IN1 = (int*)calloc(NGX, sizeof(int)); IN2 = (int*)calloc(NGX, sizeof(int)); IN3 = (int*)calloc(NGX, sizeof(int));
  IGK = (int*)calloc(NGWX*NKPT, sizeof(int));
  XKG = (double*)calloc(NGWX*NKPT, sizeof(double));
  N123 = (unsigned int*)calloc(NGWX*NKPT, sizeof(unsigned int));
  GGK = (double*)calloc(3*NGWX*NKPT, sizeof(double));
  EI1 = (double complex*)calloc((NR1+1)*NAX*NSP, sizeof(double complex));
  EI2 = (double complex*)calloc((NR2+1)*NAX*NSP, sizeof(double complex));
  EI3 = (double complex*)calloc((NR3+1)*NAX*NSP, sizeof(double complex));
  EIGR = (double complex*)calloc(NGWX*NAX*NSP*NKPT, sizeof(double complex));

  GVectorsGen();
  GkVectorsGen();
  PhaseFactor();
  g_ready = 1;
  return 0;
}

// out: NGX, NGWX, NG, NGW(1..nkpt)
void ref_sizes(int *out)
{
  out[0] = NGX; out[1] = NGWX; out[2] = NG;
  for (int ik = 0; ik < NKPT; ik++) out[3+ik] = NGW[ik];
}

void ref_scalars(double *out)
{
  out[0] = ALAT; out[1] = OMEGA; out[2] = TPIBA; out[3] = TPIBA2; out[4] = GCUT; out[5] = GCUTW;
  for (int i = 0; i < 3; i++) { out[6+i] = B1[i]; out[9+i] = B2[i]; out[12+i] = B3[i]; }
}

// in123 [3*NGX], igk/xkg/n123 [NGWX*NKPT], ggk [3*NGWX*NKPT], eigr [NGWX*NAX*NSP*NKPT]
void ref_copy(int *in123, int *igk, double *xkg, unsigned int *n123, double *ggk, double complex *eigr)
{
  for (int i = 0; i < NGX; i++) { in123[i] = IN1[i]; in123[NGX+i] = IN2[i]; in123[2*NGX+i] = IN3[i]; }
  for (int i = 0; i < NGWX*NKPT; i++) { igk[i] = IGK[i]; xkg[i] = XKG[i]; n123[i] = N123[i]; }
  for (int i = 0; i < 3*NGWX*NKPT; i++) ggk[i] = GGK[i];
  for (int i = 0; i < NGWX*NAX*NSP*NKPT; i++) eigr[i] = EIGR[i];
}

void ref_sort1(int n, double *arr, int *index) { sort1(n, arr, index); }

void ref_ortho_qr(double complex *X, int nbasis, int nstates) { ortho_qr(X, nbasis, nstates); }

void ref_eig_zheevd(double complex *A, int ldA, double *lambda, int N) { eig_zheevd(A, ldA, lambda, N); }

void ref_eig_zhegv(double complex *A, int ldA, double complex *B, int ldB, double complex *evec, int ldE,
                   int N) { eig_zhegv(A, ldA, B, ldB, evec, ldE, N); }

void ref_eig_zhegv_eval(double complex *A, int ldA, double complex *B, int ldB, double *eval,
                        double complex *evec, int ldE, int N) {
  eig_zhegv_eval(A, ldA, B, ldB, eval, evec, ldE, N);
}

void ref_random_col(double complex *A, int nelem, long a) { random_col(A, nelem, a); }

// --- energies and density mixers (xc_PZ.cpp, SimpleRhoMix.cpp, AndersonRhoMix.cpp,
// BroydenRhoMix.cpp, NormalizeRho.cpp, EvalENL.cpp) ---
void ref_xc_pz(double *rho, double *vxc, double *exc, int nnr) { xc_PZ(rho, vxc, *exc, nnr); }

void ref_simple_mix(double beta, double *rhoe, double *rho_old, int nnr, double *d) {
  SimpleRhoMix(beta, rhoe, rho_old, nnr, *d);
}

void ref_anderson_mix(int iter, double beta, double *rhoe, double *rho_old, double *f, double *g, int nnr,
                      double *d) {
  AndersonRhoMix(iter, beta, rhoe, rho_old, f, g, nnr, *d);
}

// BroydenRhoMix.cpp:63,67 take the length of its dot products from the global NNR
void ref_broyden_mix(int iter, double alpha, double w0, double *rho, double *rho_old, double *f, double *df,
                     double *u, double *a, int mixdim, int nnr, double *d) {
  NNR = nnr;
  BroydenRhoMix(iter, alpha, w0, rho, rho_old, f, df, u, a, mixdim, nnr, *d);
}

void ref_normalize_rho(double *rho, int nnr, double omega, double nel) { NormalizeRho(rho, nnr, omega, nel); }

// EvalENL.cpp:5-49 for one k-point: the globals it reads are set here
void ref_eval_enl(int nsp, const int *na, int nax, const int *l_max, const int *l_loc, int nlmax, int ngw,
                  int ngwx, int nx, double *pkg_a, double complex *eigr, double complex *c0, double *focc,
                  double wkpt, double *wnl, double complex *fnl, double *enl) {
  int ngw_[1] = {ngw};
  double wk[1] = {wkpt};
  // the set-up arrays of ref_setup stay owned by it: save and restore the globals
  const int s_nsp = NSP, s_nax = NAX, s_nlmax = NLMAX, s_ngwx = NGWX, s_nx = NX, s_nkpt = NKPT;
  int *s_na = NA, *s_lmax = L_MAX, *s_lloc = L_LOC, *s_ngw = NGW;
  double *s_pkg = PKG_A, *s_focc = FOCC, *s_wkpt = WKPT, *s_wnl = WNL;
  double complex *s_eigr = EIGR, *s_c0 = C0, *s_fnl = FNL;
  NSP = nsp; NAX = nax; NLMAX = nlmax; NGWX = ngwx; NX = nx; NKPT = 1;
  NA = const_cast<int *>(na); L_MAX = const_cast<int *>(l_max); L_LOC = const_cast<int *>(l_loc);
  NGW = ngw_; PKG_A = pkg_a; EIGR = eigr; C0 = c0; FOCC = focc; WKPT = wk; WNL = wnl; FNL = fnl;
  EvalENL();
  *enl = ENL;
  NSP = s_nsp; NAX = s_nax; NLMAX = s_nlmax; NGWX = s_ngwx; NX = s_nx; NKPT = s_nkpt;
  NA = s_na; L_MAX = s_lmax; L_LOC = s_lloc; NGW = s_ngw; PKG_A = s_pkg; FOCC = s_focc; WKPT = s_wkpt;
  WNL = s_wnl; EIGR = s_eigr; C0 = s_c0; FNL = s_fnl;
}

// --- form factors and structure factor (NLFormFactor.cpp, FormFactor.cpp,
// StructureFactor.cpp, with simpson.cpp and dsum.cpp) ---
// The radial arrays are [MAXR x nsp (x MAXAM)] as ReadPseudopotentials.cpp fills them.
void ref_nl_form_factor(int nsp, int *mmax, double *clog, double *r, double *psi, double *vion, int *l_max,
                        int *l_loc, double omega, double tpiba, int ngw, int ngwx, double *xkg, double *ggk,
                        double *xk, int nlmax, double *wnl, double *pkg, double *pkg_a) {
  int ngw_[1] = {ngw};
  const int s_nsp = NSP, s_nkpt = NKPT, s_ngwx = NGWX, s_nlmax = NLMAX;
  const double s_omega = OMEGA, s_tpiba = TPIBA;
  int *s_mmax = MMAX, *s_lmax = L_MAX, *s_lloc = L_LOC, *s_ngw = NGW;
  double *s_clog = CLOG, *s_r = R, *s_psi = PSI, *s_vion = VION, *s_xkg = XKG, *s_ggk = GGK, *s_xk = XK;
  double *s_wnl = WNL, *s_pkg = PKG, *s_pkga = PKG_A;
  NSP = nsp; NKPT = 1; NGWX = ngwx; NLMAX = nlmax; OMEGA = omega; TPIBA = tpiba;
  MMAX = mmax; L_MAX = l_max; L_LOC = l_loc; NGW = ngw_; CLOG = clog; R = r; PSI = psi; VION = vion;
  XKG = xkg; GGK = ggk; XK = xk; WNL = wnl; PKG = pkg; PKG_A = pkg_a;
  NLFormFactor();
  NSP = s_nsp; NKPT = s_nkpt; NGWX = s_ngwx; NLMAX = s_nlmax; OMEGA = s_omega; TPIBA = s_tpiba;
  MMAX = s_mmax; L_MAX = s_lmax; L_LOC = s_lloc; NGW = s_ngw; CLOG = s_clog; R = s_r; PSI = s_psi;
  VION = s_vion; XKG = s_xkg; GGK = s_ggk; XK = s_xk; WNL = s_wnl; PKG = s_pkg; PKG_A = s_pkga;
}

void ref_form_factor(int nsp, int *mmax, double *clog, double *r, double *vion, double *zv, double *rgauss,
                     int *na, int *l_loc, double omega, double tpiba2, int ng, double *g, double *vps,
                     double *rhops, double *eself) {
  const int s_nsp = NSP, s_ng = NG;
  const double s_omega = OMEGA, s_tpiba2 = TPIBA2, s_tpiba = TPIBA, s_eself = ESELF;
  int *s_mmax = MMAX, *s_na = NA, *s_lloc = L_LOC;
  double *s_clog = CLOG, *s_r = R, *s_vion = VION, *s_zv = ZV, *s_rg = RGAUSS, *s_g = G, *s_vps = VPS,
         *s_rhops = RHOPS;
  NSP = nsp; NG = ng; OMEGA = omega; TPIBA2 = tpiba2; TPIBA = sqrt(tpiba2);
  MMAX = mmax; NA = na; L_LOC = l_loc; CLOG = clog; R = r; VION = vion; ZV = zv; RGAUSS = rgauss; G = g;
  VPS = vps; RHOPS = rhops;
  FormFactor();
  *eself = ESELF;
  NSP = s_nsp; NG = s_ng; OMEGA = s_omega; TPIBA2 = s_tpiba2; TPIBA = s_tpiba; ESELF = s_eself;
  MMAX = s_mmax; NA = s_na; L_LOC = s_lloc; CLOG = s_clog; R = s_r; VION = s_vion; ZV = s_zv; RGAUSS = s_rg;
  G = s_g; VPS = s_vps; RHOPS = s_rhops;
}

// StructureFactor.cpp:5-29 on the tables ref_setup's PhaseFactor filled; sfac [nsp*NG]
void ref_structure_factor(double complex *sfac) {
  double complex *s_sfac = SFAC;
  SFAC = sfac;
  StructureFactor();
  SFAC = s_sfac;
}

void ref_free() { free_all(); }

}  // extern "C"
