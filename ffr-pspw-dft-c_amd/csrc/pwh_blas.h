// pwh_blas.h -- launchers of the small FP64 helper kernels in pwh_blas.hip.
#pragma once
#include "pwh_internal.h"

namespace pwh {
hipError_t multi_dotc(const cx<double>* V, int ld, int j, const cx<double>* f, int n,
                      cx<double>* out, hipStream_t s);
hipError_t multi_axpy(const cx<double>* V, int ld, int j, const cx<double>* coef, cx<double>* f,
                      int n, hipStream_t s);
hipError_t axpby(const cx<double>* x, double a, const cx<double>* z, double b, cx<double>* y, int n,
                 hipStream_t s);
// Lanczos steps with their scalars kept on the device: f = src - V c (c read on the device);
// y = x / sqrt(Re nrm[0]) and *tpos = sqrt(Re nrm[0]); out[i] = Re(h[i] + s[i]) (s may be null)
hipError_t multi_axpy_sub(const cx<double>* V, int ld, int j, const cx<double>* c, const cx<double>* src,
                          cx<double>* f, int n, hipStream_t s);
hipError_t scale_rnorm(const cx<double>* x, const cx<double>* nrm, cx<double>* y, int n, double* tpos,
                       hipStream_t s);
hipError_t tcol(const cx<double>* h, const cx<double>* sv, int j, double* out, hipStream_t s);
// multi_dotc over P slices of the reduction (part: j * P complex of scratch), for few
// vectors j on a long n: (P, j) blocks instead of j, partials added in a fixed order
hipError_t multi_dotc_split(const cx<double>* V, int ld, int j, const cx<double>* f, int n, cx<double>* part,
                            int P, cx<double>* out, hipStream_t s);
// p[0:n] = 0 by a kernel (a memset node does not replay reliably inside a captured graph)
hipError_t fill_zero(double* p, size_t n, hipStream_t s);
hipError_t symmetrize(cx<double>* A, int n, int ld, hipStream_t s);
// column-major n x n: the strict lower triangle <- the conjugate of the upper one
hipError_t herm_fill_lower(cx<double>* A, int n, int ld, hipStream_t s);
hipError_t zero_lower(cx<double>* A, int n, int ld, hipStream_t s);
// Column-major n x n, ld = n, one workgroup: G <- G - I, maxabs[0] = max |Re|, |Im| of it
hipError_t eye_residual(cx<double>* G, int n, double* maxabs, hipStream_t s);
// out = a in + b I (column-major n x n, ld = n; out may be in)
hipError_t axpy_eye(cx<double>* out, const cx<double>* in, double a, double b, int n, hipStream_t s);
// charge / ksum: [nb][kBandSlices] partial sums per band (sum the slices in order)
constexpr int kBandSlices = 16;
hipError_t band_sums(const cx<double>* X, int ld, int n, const double* kin, int nb, double* charge,
                     double* ksum, hipStream_t s);
// block solvers: R = lam X - Y (mode 0) or Y - X (mode 1), norms[b] = ||R_b||; part holds
// nb * kBandSlices doubles of partial sums
hipError_t resid_block(const cx<double>* X, const cx<double>* Y, const double* lam, int ld, int n, int nb,
                       int mode, cx<double>* R, double* part, double* norms, hipStream_t s);
// out[b] = Re (X_b^H Y_b)
hipError_t dotc_re(const cx<double>* X, const cx<double>* Y, int ld, int n, int nb, double* part, double* out,
                   hipStream_t s);
hipError_t scale_prec(cx<double>* R, int ld, int n, int nb, const double* res, const double* prec,
                      hipStream_t s);
hipError_t set_diag_block(cx<double>* M, int ld, int n, const double* vals, hipStream_t s);
hipError_t fill_lower(cx<double>* A, int n, int ld, hipStream_t s);
hipError_t zhegv_scale(cx<double>* V, int n, int ld, const double* b, hipStream_t s);
hipError_t resid_norms(const cx<double>* H, const cx<double>* X, const double* lam, int ld, int n,
                       int nb, double* res, hipStream_t s);
// split-K partial buffer (elements) of a projection; target: block target (0: default)
size_t proj_part_elems(int M, int N, int K, int target = 0);
// partial-sum elements the upper-triangle Gram product of N columns splits into
size_t gram_part_elems(int N, int K);
}  // namespace pwh
