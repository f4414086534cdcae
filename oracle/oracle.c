/*
 * oracle.c -- CPU FP64 restatement of the reference's Kohn-Sham H|psi> and
 * Chebyshev-filtered subspace iteration path (f-fathurrahman/ffr-pspw-dft-c).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and never by the product library.  Every function cites the reference
 * file:line whose algorithm it restates.
 *
 * Conventions kept from the reference:
 *   - arrays are 0-based here but hold the reference's values (N123 and IGK are
 *     1-based indices, exactly as GkVectorsGen.cpp:27-30 stores them);
 *   - column-major layouts: EIGR[ig + NGWX*(ia + NAX*is)], PKG_A[l + NLMAX*(ig + NGWX*is)],
 *     WNL[is + NSP*l] (common_pspw.h:17-20);
 *   - 3D grid index i1 + NR1*(i2 + NR2*i3), x fastest (fft_fftw3.cpp:29,35).
 *
 * Pinning status per function is recorded in the header comment of each
 * function ("pinned by ...").
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef double complex zc;

/* ------------------------------------------------------------------------ */
/* Lattice helpers                                                           */
/* ------------------------------------------------------------------------ */

/* determinant3x3.cpp:3-8 */
static double det3(const double *a1, const double *a2, const double *a3)
{
  return a1[0]*(a2[1]*a3[2] - a2[2]*a3[1]) -
         a1[1]*(a2[0]*a3[2] - a2[2]*a3[0]) +
         a1[2]*(a2[0]*a3[1] - a2[1]*a3[0]);
}

/* CalcRecipLatt.cpp:5-24: reciprocal vectors in units of 2*pi/ALAT. */
static void recip_latt(const double *a1, const double *a2, const double *a3,
                       double alat, double omega, double *b1, double *b2, double *b3)
{
  double s = alat/omega;
  b1[0] = (a2[1]*a3[2] - a2[2]*a3[1])*s;
  b1[1] = (a1[2]*a3[1] - a1[1]*a3[2])*s;
  b1[2] = (a1[1]*a2[2] - a1[2]*a2[1])*s;
  b2[0] = (a2[2]*a3[0] - a2[0]*a3[2])*s;
  b2[1] = (a1[0]*a3[2] - a1[2]*a3[0])*s;
  b2[2] = (a1[2]*a2[0] - a1[0]*a2[2])*s;
  b3[0] = (a2[0]*a3[1] - a2[1]*a3[0])*s;
  b3[1] = (a1[1]*a3[0] - a1[0]*a3[1])*s;
  b3[2] = (a1[0]*a2[1] - a1[1]*a2[0])*s;
}

/* matinv3x3.cpp:3-30 (column-major 3x3, same cofactor order) */
static void inv3(const double *hm, double *hi)
{
#define M(i,j) hm[((j)-1)*3 + ((i)-1)]
#define H(i,j) hi[((j)-1)*3 + ((i)-1)]
  double d11 = M(2,2)*M(3,3) - M(2,3)*M(3,2);
  double d12 = M(2,3)*M(3,1) - M(2,1)*M(3,3);
  double d13 = M(2,1)*M(3,2) - M(3,1)*M(2,2);
  double d22 = M(1,1)*M(3,3) - M(1,3)*M(3,1);
  double d23 = M(1,1)*M(3,2) - M(3,1)*M(1,2);
  double d33 = M(1,1)*M(2,2) - M(1,2)*M(2,1);
  double d21 = M(3,2)*M(1,3) - M(1,2)*M(3,3);
  double d31 = M(1,2)*M(2,3) - M(2,2)*M(1,3);
  double d32 = M(1,3)*M(2,1) - M(1,1)*M(2,3);
  double dh = M(1,1)*d11 + M(1,2)*d12 + M(1,3)*d13;
  d23 = -d23;
  H(1,1) = d11/dh; H(2,2) = d22/dh; H(3,3) = d33/dh;
  H(1,2) = d21/dh; H(1,3) = d31/dh; H(2,3) = d32/dh;
  H(2,1) = d12/dh; H(3,1) = d13/dh; H(3,2) = d23/dh;
#undef M
#undef H
}

/*
 * Fourier grid from the cutoff, Setup.cpp:20-34.  Returns the grid in nr[3].
 * Pinned by the grid sizes printed in the reference's logs.
 */
void orc_fft_grid(const double *A, double ecut, int *nr)
{
  for (int d = 0; d < 3; d++) {
    const double *a = A + 3*d;
    int n = (int)(2.0*sqrt(ecut*(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]))/M_PI);
    n = n + 2;
    if (n % 2 == 1) n = n + 1;
    nr[d] = n;
  }
}

/* Scalars of Setup.cpp:39,42,54-58: out = {ALAT, OMEGA, TPIBA, TPIBA2, GCUT, GCUTW} */
void orc_setup_scalars(const double *A, double ecut, double *out)
{
  double alat = A[0];
  double omega = det3(A, A+3, A+6);
  double tpiba = 2.*M_PI/alat;
  double tpiba2 = tpiba*tpiba;
  double gcut = 4.*ecut/tpiba/tpiba;
  out[0] = alat; out[1] = omega; out[2] = tpiba; out[3] = tpiba2;
  out[4] = gcut; out[5] = gcut/4.0;
}

/*
 * Array sizes NGX and NGWX: GVectorsGen0.cpp:5-63 followed by Setup.cpp:63-66.
 * xk: k-points [3*nkpt] in 2pi/a units.  Pinned by NGX/NGWX in the logs.
 */
void orc_gvec_sizes(const double *A, double ecut, const int *nr, const double *xk, int nkpt,
                    int *ngx_out, int *ngwx_out)
{
  double sc[6]; orc_setup_scalars(A, ecut, sc);
  double b1[3], b2[3], b3[3];
  recip_latt(A, A+3, A+6, sc[0], sc[1], b1, b2, b3);
  double gcut = sc[4], gcutw = sc[5];
  int NR1 = nr[0], NR2 = nr[1], NR3 = nr[2];
  int *ngw = (int*)malloc(sizeof(int)*(nkpt > 0 ? nkpt : 1));
  for (int ik = 0; ik < nkpt; ik++) ngw[ik] = 1;
  int jg = 2;
  for (int k = 0; k <= NR3/2; k++) {
    int n2s = (k == 0) ? 0 : -NR2/2;
    for (int j = n2s; j <= NR2/2; j++) {
      int n1s = (j == 0 && k == 0) ? 1 : -NR1/2;
      for (int i = n1s; i <= NR1/2; i++) {
        double g1 = i*b1[0] + j*b2[0] + k*b3[0];
        double g2 = i*b1[1] + j*b2[1] + k*b3[1];
        double g3 = i*b1[2] + j*b2[2] + k*b3[2];
        double g = g1*g1 + g2*g2 + g3*g3;
        for (int ik = 0; ik < nkpt; ik++) {
          const double *x = xk + 3*ik;
          double t = (g1 - x[0])*(g1 - x[0]) + (g2 - x[1])*(g2 - x[1]) + (g3 - x[2])*(g3 - x[2]);
          if (t <= gcutw) ngw[ik]++;
          t = (g1 + x[0])*(g1 + x[0]) + (g2 + x[1])*(g2 + x[1]) + (g3 + x[2])*(g3 + x[2]);
          if (t <= gcutw) ngw[ik]++;
        }
        if (g < gcut) jg++;
      }
    }
  }
  int ngx = 2*(jg - 1) + 1;
  for (int ik = 0; ik < nkpt; ik++) if (ngx < 8*ngw[ik]) ngx = 8*ngw[ik];
  int ngwx = ngx/8 + 8;
  ngx = ngwx*8 + 8;
  ngwx = ngwx + 1;
  *ngx_out = ngx; *ngwx_out = ngwx;
  free(ngw);
}

/*
 * sort1.cpp:6-55: heapsort index (FHI98MD "order"), 1-based indices, not
 * stable.  Tie order among equal |G|^2 decides the G ordering, so this is
 * restated step by step.  Pinned bit-exactly against the compiled reference.
 */
void orc_sort1(int n, const double *arr, int *index)
{
  int i, j, l, ir, indext;
  double q;
  for (i = 1; i <= n; i++) index[i-1] = i;
  if (n < 2) return;
  l = n/2 + 1;
  ir = n;
  for (;;) {
    if (l > 1) {
      l = l - 1;
      indext = index[l-1];
      q = arr[indext-1];
    } else {
      indext = index[ir-1];
      q = arr[indext-1];
      index[ir-1] = index[0];
      ir = ir - 1;
      if (ir == 1) { index[0] = indext; return; }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir && arr[index[j-1]-1] < arr[index[j]-1]) j = j + 1;
      if (q < arr[index[j-1]-1]) {
        index[i-1] = index[j-1];
        i = j;
        j = j + j;
      } else {
        j = ir + 1;
      }
    }
    index[i-1] = indext;
  }
}

/*
 * G vectors (GVectorsGen.cpp:5-137) and G+k sphere (GkVectorsGen.cpp:5-38).
 * Outputs (caller allocates):
 *   ng_out                 number of density G vectors NG
 *   in123 [3*ngx]          IN1,IN2,IN3 (signed Miller indices) for ig=1..NG
 *   ngw   [nkpt]
 *   igk   [ngwx*nkpt]      1-based index into the G list
 *   xkg   [ngwx*nkpt]      |G+k|^2 in (2pi/a)^2
 *   n123  [ngwx*nkpt]      1-based FFT index
 *   ggk   [3*ngwx*nkpt]    G (without k) in 2pi/a units, GGK(1:3,ig,ik)
 * Returns 0, or -1 on the reference's abort conditions.
 * Pinned bit-exactly against the compiled reference and by the NGW counts of its logs.
 */
int orc_gvectors(const double *A, double ecut, const int *nr, const double *xk, int nkpt,
                 int ngx, int ngwx, int *ng_out, int *in123, int *ngw, int *igk,
                 double *xkg, unsigned int *n123, double *ggk)
{
  double sc[6]; orc_setup_scalars(A, ecut, sc);
  double b1[3], b2[3], b3[3];
  recip_latt(A, A+3, A+6, sc[0], sc[1], b1, b2, b3);
  double gcut = sc[4], gcutw = sc[5];
  int NR1 = nr[0], NR2 = nr[1], NR3 = nr[2];
  int ngx21 = ngx/2 + 1;
  int *index_tmp = (int*)calloc(ngx21, sizeof(int));
  int *idum = (int*)calloc((size_t)ngx21*3, sizeof(int));
  double *dum = (double*)calloc(ngx21, sizeof(double));
  int *N1 = (int*)calloc(ngx, sizeof(int));
  int *N2 = (int*)calloc(ngx, sizeof(int));
  int *N3 = (int*)calloc(ngx, sizeof(int));
  double *GG = (double*)calloc((size_t)(ngx+1)*3, sizeof(double));
  int *IN1 = in123, *IN2 = in123 + ngx, *IN3 = in123 + 2*ngx;
  int rc = 0;

  int jg = 2;
  for (int k = 0; k <= NR3/2; k++) {
    int n2s = (k == 0) ? 0 : -NR2/2;
    for (int j = n2s; j <= NR2/2; j++) {
      int n1s = (j == 0 && k == 0) ? 1 : -NR1/2;
      for (int i = n1s; i <= NR1/2; i++) {
        double g1 = i*b1[0] + j*b2[0] + k*b3[0];
        double g2 = i*b1[1] + j*b2[1] + k*b3[1];
        double g3 = i*b1[2] + j*b2[2] + k*b3[2];
        double g = g1*g1 + g2*g2 + g3*g3;
        if (g <= gcut) {
          if (jg > ngx21) { rc = -1; goto done; }
          dum[jg-1] = g;
          idum[(jg-1)] = i;
          idum[(jg-1) + ngx21] = j;
          idum[(jg-1) + 2*ngx21] = k;
          jg++;
        }
      }
    }
  }
  int ng_stop = jg - 1;
  orc_sort1(ng_stop, dum, index_tmp);
  if (dum[index_tmp[0]-1] != 0 || index_tmp[0] != 1) { rc = -1; goto done; }

  IN1[0] = 0; IN2[0] = 0; IN3[0] = 0;
  GG[0] = 0.0; GG[ngx+1] = 0.0; GG[2*(ngx+1)] = 0.0;
  N1[0] = 1; N2[0] = 1; N3[0] = 1;
  int NG = 2;
  for (jg = 2; jg <= ng_stop; jg++) {
    if (NG + 1 > ngx) { rc = -1; goto done; }
    int id = index_tmp[jg-1];
    int i = idum[id-1], j = idum[id-1 + ngx21], k = idum[id-1 + 2*ngx21];
    IN1[NG-1] = i; IN2[NG-1] = j; IN3[NG-1] = k;
    IN1[NG] = -i; IN2[NG] = -j; IN3[NG] = -k;
    GG[(NG-1)]              = i*b1[0] + j*b2[0] + k*b3[0];
    GG[(NG-1) + (ngx+1)]    = i*b1[1] + j*b2[1] + k*b3[1];
    GG[(NG-1) + 2*(ngx+1)]  = i*b1[2] + j*b2[2] + k*b3[2];
    GG[NG]                  = -GG[NG-1];
    GG[NG + (ngx+1)]        = -GG[(NG-1) + (ngx+1)];
    GG[NG + 2*(ngx+1)]      = -GG[(NG-1) + 2*(ngx+1)];
    N1[NG-1] = i + 1; if (i < 0) N1[NG-1] += NR1;
    N2[NG-1] = j + 1; if (j < 0) N2[NG-1] += NR2;
    N3[NG-1] = k + 1; if (k < 0) N3[NG-1] += NR3;
    i = -i; j = -j; k = -k;
    N1[NG] = i + 1; if (i < 0) N1[NG] += NR1;
    N2[NG] = j + 1; if (j < 0) N2[NG] += NR2;
    N3[NG] = k + 1; if (k < 0) N3[NG] += NR3;
    NG += 2;
  }
  NG = NG - 1;
  *ng_out = NG;

  for (int ik = 0; ik < nkpt; ik++) {
    const double *x = xk + 3*ik;
    ngw[ik] = 0;
    for (int ig = 1; ig <= NG; ig++) {
      double gx = GG[ig-1] + x[0];
      double gy = GG[(ig-1) + (ngx+1)] + x[1];
      double gz = GG[(ig-1) + 2*(ngx+1)] + x[2];
      double t = gx*gx + gy*gy + gz*gz;
      if (t <= gcutw) {
        ngw[ik]++;
        int igp = ngw[ik];
        if (igp > ngwx) { rc = -1; goto done; }
        size_t o = (size_t)(igp-1) + (size_t)ngwx*ik;
        igk[o] = ig;
        xkg[o] = t;
        n123[o] = (unsigned int)(N1[ig-1] + NR1*((N2[ig-1]-1) + NR2*(N3[ig-1]-1)));
        ggk[3*o + 0] = GG[ig-1];
        ggk[3*o + 1] = GG[(ig-1) + (ngx+1)];
        ggk[3*o + 2] = GG[(ig-1) + 2*(ngx+1)];
      }
    }
  }
done:
  free(index_tmp); free(idum); free(dum); free(N1); free(N2); free(N3); free(GG);
  return rc;
}

/*
 * Structure-factor phases EIGR(G,a) = exp(-i G.tau_a): PhaseFactor.cpp:5-80.
 * tau [3*nax*nsp] Cartesian bohr (TAU(1:3,ia,is)), na[nsp], in123 as returned
 * by orc_gvectors (stride ngx), eigr [ngwx*nax*nsp*nkpt].
 * Pinned bit-exactly against the compiled reference.
 */
void orc_phase_factor(const double *A, const int *nr, int nsp, const int *na, int nax,
                      const double *tau, int nkpt, const int *ngw, int ngwx, const int *igk,
                      int ngx, const int *in123, zc *eigr)
{
  int NR1 = nr[0], NR2 = nr[1], NR3 = nr[2];
  int iz1 = (NR1+1)/2 + 1, iz2 = (NR2+1)/2 + 1, iz3 = (NR3+1)/2 + 1;
  int NR11 = NR1 + 1, NR21 = NR2 + 1, NR31 = NR3 + 1;
  double xmat[9], xmati[9];
  for (int i = 0; i < 3; i++) { xmat[i] = A[i]; xmat[3+i] = A[3+i]; xmat[6+i] = A[6+i]; }
  inv3(xmat, xmati);
  zc *EI1 = (zc*)calloc((size_t)NR11*nax*nsp, sizeof(zc));
  zc *EI2 = (zc*)calloc((size_t)NR21*nax*nsp, sizeof(zc));
  zc *EI3 = (zc*)calloc((size_t)NR31*nax*nsp, sizeof(zc));
  for (int is = 0; is < nsp; is++) {
    for (int ia = 0; ia < na[is]; ia++) {
      const double *t = tau + 3*(ia + nax*is);
      double taup[3];
      for (int i = 0; i < 3; i++) {
        double s1 = xmati[i + 0]*t[0];
        s1 = s1 + xmati[i + 3]*t[1];
        s1 = s1 + xmati[i + 6]*t[2];
        taup[i] = s1;
      }
      zc *e1 = EI1 + (size_t)NR11*(ia + nax*is);
      zc *e2 = EI2 + (size_t)NR21*(ia + nax*is);
      zc *e3 = EI3 + (size_t)NR31*(ia + nax*is);
      e1[iz1-1] = 1.0; e2[iz2-1] = 1.0; e3[iz3-1] = 1.0;
      double ar1 = -2.0*M_PI*taup[0], ar2 = -2.0*M_PI*taup[1], ar3 = -2.0*M_PI*taup[2];
      for (int i = 1; i <= (NR1+1)/2; i++) {
        double c = cos(i*ar1), s = sin(i*ar1);
        if (iz1+i-1 <= NR1) e1[iz1+i-1] = CMPLX(c, s); e1[iz1-i-1] = CMPLX(c, -s);
      }
      for (int j = 1; j <= (NR2+1)/2; j++) {
        double c = cos(j*ar2), s = sin(j*ar2);
        if (iz2+j-1 <= NR2) e2[iz2+j-1] = CMPLX(c, s); e2[iz2-j-1] = CMPLX(c, -s);
      }
      for (int k = 1; k <= (NR3+1)/2; k++) {
        double c = cos(k*ar3), s = sin(k*ar3);
        if (iz3+k-1 <= NR3) e3[iz3+k-1] = CMPLX(c, s); e3[iz3-k-1] = CMPLX(c, -s);
      }
    }
  }
  const int *IN1 = in123, *IN2 = in123 + ngx, *IN3 = in123 + 2*ngx;
  for (int ik = 0; ik < nkpt; ik++)
    for (int is = 0; is < nsp; is++)
      for (int ia = 0; ia < nax; ia++)
        for (int ig = 0; ig < ngw[ik]; ig++) {
          int igp = igk[ig + ngwx*ik];
          zc a = EI1[(iz1 + IN1[igp-1] - 1) + (size_t)NR11*(ia + nax*is)];
          zc b = EI2[(iz2 + IN2[igp-1] - 1) + (size_t)NR21*(ia + nax*is)];
          zc c = EI3[(iz3 + IN3[igp-1] - 1) + (size_t)NR31*(ia + nax*is)];
          /* left-to-right complex products, textbook formula */
          double abr = creal(a)*creal(b) - cimag(a)*cimag(b);
          double abi = creal(a)*cimag(b) + cimag(a)*creal(b);
          double r = abr*creal(c) - abi*cimag(c);
          double im = abr*cimag(c) + abi*creal(c);
          eigr[ig + (size_t)ngwx*(ia + (size_t)nax*(is + (size_t)nsp*ik))] = CMPLX(r, im);
        }
  free(EI1); free(EI2); free(EI3);
}

/* ------------------------------------------------------------------------ */
/* FFT: fft_fftw3.cpp:13-46 conventions.                                     */
/* ------------------------------------------------------------------------ */

/*
 * 1D complex DFT of length n (any n), recursive mixed radix on the smallest
 * prime factor, out of place: out[k] = sum_j in[j*stride] w^{jk}, with
 * w = exp(sign*2*pi*i/n).  tw holds exp(sign*2*pi*i*m/N) for the top-level N,
 * tstep = N/n.
 */
static void dft_rec(int n, const zc *in, int stride, zc *out, const zc *tw, int tstep, zc *scratch)
{
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (p*p <= n && n % p) p++;
  if (n % p) p = n;
  int m = n/p;
  /* sub-transforms: out[q*m + k] = DFT_m of in[(q + p*j)*stride] */
  for (int q = 0; q < p; q++)
    dft_rec(m, in + q*stride, stride*p, out + q*m, tw, tstep*p, scratch);
  /* butterflies of radix p */
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      /* twiddle w_n^{q*k} */
      scratch[q] = out[q*m + k]*tw[((size_t)q*k*tstep) % ((size_t)n*tstep)];
    }
    for (int r = 0; r < p; r++) {
      zc s = 0;
      for (int q = 0; q < p; q++)
        s += scratch[q]*tw[((size_t)q*r*m*tstep) % ((size_t)n*tstep)];
      out[r*m + k] = s;
    }
    /* out[r*m + k] written after all reads of column k; the q-loop only
       reads index q*m + k, all captured in scratch first. */
  }
}

static void fft_lines(zc *data, int n, int nlines, int line_stride, int elem_stride,
                      int sign, const int *line_offsets)
{
  zc *tw = (zc*)malloc(sizeof(zc)*n);
  for (int m = 0; m < n; m++) {
    double a = sign*2.0*M_PI*(double)m/(double)n;
    tw[m] = CMPLX(cos(a), sin(a));
  }
  zc *buf = (zc*)malloc(sizeof(zc)*n);
  zc *out = (zc*)malloc(sizeof(zc)*n);
  zc *scr = (zc*)malloc(sizeof(zc)*n);
  for (int l = 0; l < nlines; l++) {
    zc *p = data + (line_offsets ? line_offsets[l] : (size_t)l*line_stride);
    for (int j = 0; j < n; j++) buf[j] = p[(size_t)j*elem_stride];
    dft_rec(n, buf, 1, out, tw, 1, scr);
    for (int j = 0; j < n; j++) p[(size_t)j*elem_stride] = out[j];
  }
  free(tw); free(buf); free(out); free(scr);
}

/*
 * In-place 3D DFT, fft_fftw3.cpp:13-46: inverse -> exp(+i..) unscaled
 * (FFTW_BACKWARD); forward -> exp(-i..) then scaled by 1/N.  Data layout x
 * fastest: data[i1 + NR1*(i2 + NR2*i3)].
 * Pinned against numpy.fft, a brute-force DFT and the sandbox known answer.
 */
void orc_fft3d(zc *data, int nr1, int nr2, int nr3, int inverse)
{
  int sign = inverse ? +1 : -1;
  int n = nr1*nr2*nr3;
  /* x lines */
  fft_lines(data, nr1, nr2*nr3, nr1, 1, sign, NULL);
  /* y lines: for each (i1, i3) */
  int *off = (int*)malloc(sizeof(int)*(size_t)nr1*nr3);
  for (int i3 = 0; i3 < nr3; i3++)
    for (int i1 = 0; i1 < nr1; i1++) off[i1 + nr1*i3] = i1 + nr1*nr2*i3;
  fft_lines(data, nr2, nr1*nr3, 0, nr1, sign, off);
  free(off);
  /* z lines */
  fft_lines(data, nr3, nr1*nr2, 1, nr1*nr2, sign, NULL);
  if (!inverse) {
    double s = 1.0/(double)n;
    for (int i = 0; i < n; i++) data[i] *= s;
  }
}

/* ------------------------------------------------------------------------ */
/* H|psi>                                                                    */
/* ------------------------------------------------------------------------ */

/* The reference's globals read by ApplyHam_block (common_pspw.h:43-148). */
typedef struct {
  int nr1, nr2, nr3, nnr;
  int ngwx;
  const unsigned int *n123;   /* [ngwx*nkpt] 1-based */
  const double *xkg;          /* [ngwx*nkpt] */
  double tpiba2;
  int nsp, nax, nlmax;
  const int *na;              /* [nsp] */
  const int *l_max, *l_loc;   /* [nsp] */
  const double *wnl;          /* [nsp*nlmax], WNL(is,l) at is + nsp*l */
  const zc *eigr;             /* [ngwx*nax*nsp*nkpt] */
  const double *pkg_a;        /* [nlmax*ngwx*nsp*nkpt] */
} orc_sys;

/*
 * ApplyHam_block.cpp:5-172 (and ApplyHam.cpp:5-142 for nstates = 1):
 * vout(:,ist) = 0.5*XKG*TPIBA2*vin + FFT_fwd[rho * FFT_inv[vin]] + nonlocal.
 * vin/vout [ngw_ik*nstates] with ld = ngw_ik; ik is 1-based; rho = V_eff [nnr].
 * Returns -1 where the reference aborts (nnr != NNR, lm_end > 4).
 * Pinned by: V=const and kinetic closed forms, a dense plane-wave H built
 * from brute-force DFT matrix elements, Hermiticity, and the dense beta D beta^H.
 */
int orc_apply_ham_block(const orc_sys *S, const zc *vin, zc *vout, int ngw_ik, int ik,
                        const double *rho, int nnr, int nstates)
{
  if (nnr != S->nnr) return -1;
  int NGWX = S->ngwx, NSP = S->nsp, NAX = S->nax, NLMAX = S->nlmax;
  zc *c_fft = (zc*)malloc(sizeof(zc)*(size_t)S->nnr);
  zc *t = (zc*)malloc(sizeof(zc)*(size_t)NGWX*NLMAX);
  const unsigned int *n123 = S->n123 + (size_t)NGWX*(ik-1);
  const double *xkg = S->xkg + (size_t)NGWX*(ik-1);
  int rc = 0;
  for (int ist = 0; ist < nstates; ist++) {
    const zc *vi = vin + (size_t)ngw_ik*ist;
    zc *vo = vout + (size_t)ngw_ik*ist;
    memset(c_fft, 0, sizeof(zc)*(size_t)S->nnr);
    for (int ig = 0; ig < ngw_ik; ig++) c_fft[n123[ig]-1] = vi[ig];
    orc_fft3d(c_fft, S->nr1, S->nr2, S->nr3, 1);
    for (int i = 0; i < S->nnr; i++) c_fft[i] = rho[i]*c_fft[i];
    orc_fft3d(c_fft, S->nr1, S->nr2, S->nr3, 0);
    for (int ig = 0; ig < ngw_ik; ig++)
      vo[ig] = 0.5*xkg[ig]*S->tpiba2*vi[ig] + c_fft[n123[ig]-1];

    for (int is = 0; is < NSP; is++) {
      int lm_end = S->l_max[is]*S->l_max[is] - 2*S->l_loc[is] + 1;
      if (lm_end > 4) { rc = -1; goto out; }
      double rr[4] = {0, 0, 0, 0};
      for (int il = 0; il < lm_end; il++) rr[il] = S->wnl[is + NSP*il];
      for (int ia = 0; ia < S->na[is]; ia++) {
        const zc *eg = S->eigr + (size_t)NGWX*(ia + (size_t)NAX*(is + (size_t)NSP*(ik-1)));
        const double *pk = S->pkg_a + (size_t)NLMAX*NGWX*(is + (size_t)NSP*(ik-1));
        /* t = conj-phase * PKG_A, ApplyHam_block.cpp:88-95 */
        for (int il = 0; il < lm_end; il++)
          for (int ig = 0; ig < ngw_ik; ig++) {
            double p = pk[il + (size_t)NLMAX*ig];
            t[ig + (size_t)NGWX*il] = CMPLX(creal(eg[ig])*p, cimag(eg[ig])*(-p));
          }
        /* zdotu + accumulate, ApplyHam_block.cpp:97-163 */
        zc tt[4] = {0, 0, 0, 0};
        for (int il = 0; il < lm_end; il++) {
          zc s = 0;
          for (int ig = 0; ig < ngw_ik; ig++) s += vi[ig]*t[ig + (size_t)NGWX*il];
          tt[il] = rr[il]*s;
        }
        for (int ig = 0; ig < ngw_ik; ig++) {
          zc acc = 0;
          for (int il = 0; il < lm_end; il++) acc += pk[il + (size_t)NLMAX*ig]*tt[il];
          vo[ig] = vo[ig] + eg[ig]*acc;
        }
      }
    }
  }
out:
  free(c_fft); free(t);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Chebyshev filter and Lanczos                                              */
/* ------------------------------------------------------------------------ */

/*
 * chebyfilt.cpp:6-74: degree-m scaled Chebyshev filter applied in place to
 * X [ngw*nstates].  The reference hard-wires nstates = NX+2 (chebyfilt.cpp:20);
 * here it is an argument.  Same sigma recurrence and operation order
 * (zaxpy/zdscal/zcopy sequence).
 * Pinned by the closed form T_m((H-c)/e) X / T_m(-2) for a diagonal H.
 */
int orc_chebyfilt(const orc_sys *S, int ik, zc *X, int ngw, int nstates, int degree,
                  double lb, double ub, const double *rho, int nnr)
{
  double e = (ub - lb)/2.0;
  double c = (ub + lb)/2.0;
  double sigma = e/(lb - ub);
  double sigma1 = sigma, sigma2;
  size_t nelem = (size_t)ngw*nstates;
  zc *Y = (zc*)malloc(sizeof(zc)*nelem);
  zc *Y1 = (zc*)malloc(sizeof(zc)*nelem);
  int rc = orc_apply_ham_block(S, X, Y, ngw, ik, rho, nnr, nstates);
  if (rc) goto out;
  for (size_t i = 0; i < nelem; i++) Y[i] = Y[i] + (-c)*X[i];
  {
    double scale = sigma1/e;
    for (size_t i = 0; i < nelem; i++) Y[i] = scale*Y[i];
  }
  for (int id = 2; id <= degree; id++) {
    sigma2 = 1.0/(2.0/sigma1 - sigma);
    rc = orc_apply_ham_block(S, Y, Y1, ngw, ik, rho, nnr, nstates);
    if (rc) goto out;
    double scale = 2.0*sigma2/e;
    for (size_t i = 0; i < nelem; i++) Y1[i] = scale*Y1[i];
    double za = -c*2.0*sigma2/e;
    for (size_t i = 0; i < nelem; i++) Y1[i] = Y1[i] + za*Y[i];
    za = -sigma*sigma2;
    for (size_t i = 0; i < nelem; i++) Y1[i] = Y1[i] + za*X[i];
    memcpy(X, Y, sizeof(zc)*nelem);
    memcpy(Y, Y1, sizeof(zc)*nelem);
    sigma = sigma2;
  }
  memcpy(X, Y, sizeof(zc)*nelem);
out:
  free(Y); free(Y1);
  return rc;
}

static zc zdotc(int n, const zc *x, const zc *y)
{
  zc s = 0;
  for (int i = 0; i < n; i++) s += conj(x[i])*y[i];
  return s;
}

/*
 * lanczos.cpp:8-101: k-step Lanczos with one re-orthogonalisation pass.
 * v0 [ngw] is normalised in place (as the reference does); T [ldT*k]
 * column-major gets the upper triangle of the projected matrix plus the
 * subdiagonal norms; f [ngw] is the final residual.  Vout [ngw*k] optional.
 * Pinned by the Lanczos relation H V = V T + f e_k^T and V^H V = I.
 */
int orc_lanczos(const orc_sys *S, int ik, zc *v0, int ngw, int k, double *T, int ldT, zc *f,
                const double *rho, int nnr, zc *Vout)
{
  zc *V = (zc*)malloc(sizeof(zc)*(size_t)ngw*k);
  zc *Hv = (zc*)malloc(sizeof(zc)*(size_t)ngw);
  zc *h = (zc*)malloc(sizeof(zc)*k);
  zc *s = (zc*)malloc(sizeof(zc)*k);
  int rc = 0;
  for (int i = 0; i < ldT*k; i++) T[i] = 0.0;
  zc tz = zdotc(ngw, v0, v0);
  double beta = 1.0/sqrt(creal(tz));
  for (int i = 0; i < ngw; i++) v0[i] = beta*v0[i];
  memcpy(V, v0, sizeof(zc)*ngw);
  rc = orc_apply_ham_block(S, V, Hv, ngw, ik, rho, nnr, 1);
  if (rc) goto out;
  tz = zdotc(ngw, V, Hv);
  h[0] = tz;
  T[0] = creal(tz);
  for (int ig = 0; ig < ngw; ig++) f[ig] = Hv[ig] - h[0]*V[ig];
  s[0] = zdotc(ngw, V, f);
  h[0] = h[0] + s[0];
  for (int ig = 0; ig < ngw; ig++) f[ig] = f[ig] + (-s[0])*V[ig];
  for (int j = 2; j <= k; j++) {
    tz = zdotc(ngw, f, f);
    T[(j-1) + (size_t)ldT*(j-2)] = sqrt(creal(tz));
    beta = 1.0/sqrt(creal(tz));
    zc *vj = V + (size_t)ngw*(j-1);
    for (int ig = 0; ig < ngw; ig++) vj[ig] = beta*f[ig];
    rc = orc_apply_ham_block(S, vj, Hv, ngw, ik, rho, nnr, 1);
    if (rc) goto out;
    memcpy(f, Hv, sizeof(zc)*ngw);
    for (int i = 1; i <= j; i++) {
      h[i-1] = zdotc(ngw, V + (size_t)ngw*(i-1), Hv);
      zc za = -h[i-1];
      const zc *vi = V + (size_t)ngw*(i-1);
      for (int ig = 0; ig < ngw; ig++) f[ig] = f[ig] + za*vi[ig];
    }
    for (int i = 1; i <= j; i++) s[i-1] = zdotc(ngw, V + (size_t)ngw*(i-1), f);
    for (int i = 1; i <= j; i++) h[i-1] = h[i-1] + s[i-1];
    for (int i = 1; i <= j; i++) {
      zc za = -s[i-1];
      const zc *vi = V + (size_t)ngw*(i-1);
      for (int ig = 0; ig < ngw; ig++) f[ig] = f[ig] + za*vi[ig];
    }
    for (int i = 1; i <= j; i++) T[(i-1) + (size_t)ldT*(j-1)] = creal(h[i-1]);
  }
  if (Vout) memcpy(Vout, V, sizeof(zc)*(size_t)ngw*k);
out:
  free(V); free(Hv); free(h); free(s);
  return rc;
}

/*
 * induced_norm_mat, ChebySCFLoop.cpp:7-34: || A * (1/sqrt(N)) 1 ||_2^2
 * (the reference returns the squared 2-norm).
 */
double orc_induced_norm_mat(const double *A, int ldA, int N)
{
  double d = 1.0/sqrt((double)N);
  double res = 0.0;
  for (int i = 0; i < N; i++) {
    double v = 0.0;
    for (int j = 0; j < N; j++) v += A[i + (size_t)ldA*j]*d;
    res += v*v;
  }
  return res;
}

/*
 * Same as orc_apply_ham_block, with the band loop split over nthreads OpenMP
 * threads (bands are independent in ApplyHam_block.cpp:28-167).  Used only as
 * the multi-core CPU baseline.
 */
int orc_apply_ham_block_par(const orc_sys *S, const zc *vin, zc *vout, int ngw_ik, int ik,
                            const double *rho, int nnr, int nstates, int nthreads)
{
  int rc = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(|:rc)
  for (int ist = 0; ist < nstates; ist++)
    rc |= orc_apply_ham_block(S, vin + (size_t)ngw_ik*ist, vout + (size_t)ngw_ik*ist, ngw_ik, ik,
                              rho, nnr, 1);
  return rc;
}

/*
 * EvalRhoPsi.cpp:83-157 (the ldEigvec form; the NGWX form at :5-79 is the same with
 * ld = NGWX).  For each of nstates states: scatter coefficients ig = 1 .. NGW_ik-1 (the
 * reference's loop stops one short of NGW_ik, :117) onto the grid through
 * IGK -> (N1,N2,N3), which equals N123 - 1 (GkVectorsGen.cpp:30); charge ss = sum |c|^2
 * and sk = sum XKG |c|^2 over the same range; inverse FFT (unscaled, e^{+i});
 * rhoe += |c_fft|^2 * w / omega, with w = WKPT(ik) * FOCC(ist, ik) passed per state.
 * ekin_out = sum_states sk * w / 2 * TPIBA2 (what the reference adds to EKIN, :154).
 * charge[ist] = ss; returns the 0-based index of the first state with |ss - 1| > 6e-3
 * (where the reference prints and aborts, :129-133), or -1.  That state's density is
 * not added, as in the reference.
 * Pinned by: a single plane wave (|psi(r)|^2 constant), two plane waves (closed-form
 * cosine), Parseval (sum_r |psi|^2 = NNR sum_G |c|^2), the skipped last coefficient.
 */
int orc_eval_rho_psi(const orc_sys *S, int ik, const zc *eigvec, int ld, int ngw_ik, int nstates,
                     const double *w, double omega, double *rhoe, double *ekin_out, double *charge)
{
  const unsigned int *n123 = S->n123 + (size_t)S->ngwx*(ik-1);
  const double *xkg = S->xkg + (size_t)S->ngwx*(ik-1);
  zc *c_fft = (zc*)malloc(sizeof(zc)*(size_t)S->nnr);
  double xkin = 0.0;
  int bad = -1;
  for (int ist = 0; ist < nstates; ist++) {
    const zc *c = eigvec + (size_t)ld*ist;
    const double wdotfbo = w[ist] / omega, wdot2fb = w[ist] * 0.5;
    double ss = 0.0, sk = 0.0;
    memset(c_fft, 0, sizeof(zc)*(size_t)S->nnr);
    for (int ig = 0; ig < ngw_ik - 1; ig++) {
      c_fft[n123[ig]-1] = c[ig];
      const double a = cabs(c[ig]);
      const double scg = a*a;
      ss += scg;
      sk += xkg[ig]*scg;
    }
    charge[ist] = ss;
    if (fabs(ss - 1.0) > 6.e-3 && bad < 0) { bad = ist; break; }
    xkin += sk*wdot2fb;
    orc_fft3d(c_fft, S->nr1, S->nr2, S->nr3, 1);
    for (int i = 0; i < S->nnr; i++) {
      const double a = cabs(c_fft[i]);
      rhoe[i] += a*a*wdotfbo;
    }
  }
  *ekin_out = xkin*S->tpiba2;
  free(c_fft);
  return bad;
}
