// pwh_scf.h -- launchers of the SCF-step kernels in pwh_scf.hip (energies, xc_PZ,
// density mixing, NormalizeRho, the EvalENL sums).
#pragma once
#include "pwh_internal.h"

namespace pwh {

constexpr int kScfBlocks = 256;    // fixed grid of the reductions: one partial per block
constexpr int kScfThreads = 256;
constexpr int kMaxTerms = 6;

// out (and out2 when set) = sum_k w[k] v[k]; with p set, the kernel's partials hold
// sum (p - q)^2 (q may be null)
struct Comb {
  int nt;
  double w[kMaxTerms];
  const double* v[kMaxTerms];
  double* out;
  double* out2;
  const double* p;
  const double* q;
};

hipError_t xc_pz(const double* rho, size_t n, double* vxc, double* eps, unsigned char* neg,
                 const cx<double>* c, double* out, double* part, hipStream_t s);
hipError_t xc_pz_fix(const double* rho, size_t n, const unsigned char* neg, double* vxc, double* eps,
                     const cx<double>* c, double* out, hipStream_t s);
hipError_t dot_partials(const double* x, const double* y, size_t n, double* part, hipStream_t s);
hipError_t energy_g(cx<double>* c, const int* gidx, const double* g2, const cx<double>* vp,
                    const cx<double>* rp, int ng, double tpiba2, double* part, hipStream_t s);
hipError_t real_to_cx(const double* a, const double* b, size_t n, cx<double>* c, hipStream_t s);
hipError_t kerker_filter(cx<double>* c, const int* idx, const double* xkg, int ngw, double beta, double q0,
                         hipStream_t s);
hipError_t cx_real(const cx<double>* c, size_t n, double* out, hipStream_t s);
hipError_t comb(const Comb& cb, size_t n, double* part, hipStream_t s);
hipError_t clip_sum(double* rho, size_t n, double* part, hipStream_t s);
hipError_t scale_vec(double* x, size_t n, double num, double den, hipStream_t s);
hipError_t to_float(const double* x, size_t n, float* y, hipStream_t s);
hipError_t enl_rows(const cx<double>* C, int ld, int rows, int nb, const double* focc, double* out,
                    hipStream_t s);

}  // namespace pwh
