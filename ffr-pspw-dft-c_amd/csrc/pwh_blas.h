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
hipError_t symmetrize(cx<double>* A, int n, int ld, hipStream_t s);
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
}  // namespace pwh
