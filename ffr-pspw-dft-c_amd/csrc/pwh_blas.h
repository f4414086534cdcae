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
// charge / ksum: [nb][kBandSlices] partial sums per band (sum the slices in order)
constexpr int kBandSlices = 16;
hipError_t band_sums(const cx<double>* X, int ld, int n, const double* kin, int nb, double* charge,
                     double* ksum, hipStream_t s);
hipError_t resid_norms(const cx<double>* H, const cx<double>* X, const double* lam, int ld, int n,
                       int nb, double* res, hipStream_t s);
size_t proj_part_elems(int M, int N, int K);
}  // namespace pwh
