// pwh_scf.hip -- kernels of the SCF steps around the eigensolver: the local energy terms
// and effective potential (EvalEnergy.cpp), the Perdew-Zunger LDA (xc_PZ.cpp), the
// density mixers (SimpleRhoMix.cpp, KerkerRhoMix.cpp, AndersonRhoMix.cpp, PulayRhoMix.cpp,
// BroydenRhoMix.cpp), NormalizeRho.cpp and the nonlocal energy sum of EvalENL.cpp.
//
// All of it is FP64 and HBM-bound pointwise work over the NNR grid (or the NG density
// G-vectors).  Every reduction is written as one partial per block of a fixed-size grid
// (kScfBlocks blocks, each a fixed-order tree), and the host adds the partials in order,
// so results are bitwise repeatable.
#include "pwh_scf.h"

namespace pwh {

namespace {

__device__ __forceinline__ void block_sum(double v, double* sh, double* out) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int h = kScfThreads / 2; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

// xc_PZ.cpp:9-62, one point: returns ux + uc, writes ex + ec
__device__ __forceinline__ double pz_point(double r, double* e) {
  const double A = -0.1423, B = 1.0529, C = 0.3334;
  const double a = 0.0311, b = -0.0480, c = 0.0020, d = -0.0116;
  if (r <= 1.e-13) {
    *e = 0.0;
    return 0.0;
  }
  const double pi = 3.14159265358979323846;
  const double ax = -(3.0 / 4.0) * pow(3.0 / 2.0 / pi, 2.0 / 3.0);
  const double rs = pow(3.0 / (4.0 * pi * r), 1.0 / 3.0);
  const double rrs = sqrt(rs);
  const double ex = ax / rs;
  const double ux = 4.0 / 3.0 * ex;
  double ec, uc;
  if (rs >= 1.0) {
    const double bpr = 1.0 + B * rrs + C * rs;
    ec = A / bpr;
    uc = (1.0 + 7.0 / 6.0 * B * rrs + 4.0 / 3.0 * C * rs) * ec / bpr;
  } else {
    const double rsl = log(rs);
    ec = a * rsl + b + c * rs * rsl + d * rs;
    uc = (2.0 * d - c) / 3.0 * rs + b - a / 3.0 + a * rsl + 2.0 * c / 3.0 * rs * rsl;
  }
  *e = ex + ec;
  return ux + uc;
}

constexpr double kNegRho = -0.1;   // xc_PZ.cpp:26

}  // namespace

// vxc, eps = ex + ec per point; points below -0.1 are marked (neg = 1) and counted;
// with c, out = Re(c) + vxc (EvalEnergy.cpp:131-134)
__global__ __launch_bounds__(kScfThreads) void k_xc(const double* __restrict__ rho, size_t n,
                                                    double* __restrict__ vxc, double* __restrict__ eps,
                                                    unsigned char* __restrict__ neg,
                                                    const cx<double>* __restrict__ c, double* __restrict__ out,
                                                    double* __restrict__ part) {
  __shared__ double sh[kScfThreads];
  double cnt = 0.0;
  for (size_t i = (size_t)blockIdx.x * kScfThreads + threadIdx.x; i < n; i += (size_t)kScfBlocks * kScfThreads) {
    const double r = rho[i];
    double v = 0.0, e = 0.0;
    const bool bad = r < kNegRho;
    if (!bad) v = pz_point(r, &e);
    else cnt += 1.0;
    neg[i] = bad ? 1 : 0;
    vxc[i] = v;
    eps[i] = e;
    if (c) out[i] = c[i].x + v;
  }
  block_sum(cnt, sh, part + blockIdx.x);
}

// Points below -0.1 take the values of the nearest earlier point that is not (the
// reference's loop-carried locals); none before: zero.  Runs only when some exist.
__global__ void k_xc_fix(const double* __restrict__ rho, size_t n, const unsigned char* __restrict__ neg,
                         double* __restrict__ vxc, double* __restrict__ eps, const cx<double>* __restrict__ c,
                         double* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (!neg[i]) continue;
    double v = 0.0, e = 0.0;
    for (size_t j = i; j-- > 0;) {
      if (!neg[j]) {
        v = vxc[j];
        e = eps[j];
        break;
      }
    }
    vxc[i] = v;
    eps[i] = e;
    if (c) out[i] = c[i].x + v;
  }
}

// part[block] = sum x * y over the block's points
__global__ __launch_bounds__(kScfThreads) void k_dot(const double* __restrict__ x, const double* __restrict__ y,
                                                     size_t n, double* __restrict__ part) {
  __shared__ double sh[kScfThreads];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kScfThreads + threadIdx.x; i < n; i += (size_t)kScfBlocks * kScfThreads)
    s += x[i] * y[i];
  block_sum(s, sh, part + blockIdx.x);
}

// EvalEnergy.cpp:57-107 over the density G-vectors: c holds the forward transform of the
// density; each G reads its coefficient and replaces it by the electrostatic potential
// (Hartree + local pseudopotential).  Partials: Re eh, Re eps, eh2 (electrons only),
// eg2 (pseudo-charge only), one set per block.
__global__ __launch_bounds__(kScfThreads) void k_energy_g(cx<double>* __restrict__ c, const int* __restrict__ gidx,
                                                          const double* __restrict__ g2,
                                                          const cx<double>* __restrict__ vp,
                                                          const cx<double>* __restrict__ rp, int ng, double tpiba2,
                                                          double* __restrict__ part) {
  __shared__ double sh[kScfThreads];
  const double fourpi = 4.0 * 3.14159265358979323846;
  double eh = 0.0, ep = 0.0, e2 = 0.0, g2p = 0.0;
  for (int ig = blockIdx.x * kScfThreads + threadIdx.x; ig < ng; ig += kScfBlocks * kScfThreads) {
    const int k = gidx[ig];
    const cx<double> rhet = c[k], v = vp[ig];
    if (ig == 0) {
      // G = 0: only the pseudopotential energy of the mean density (:55-57)
      ep += v.x * rhet.x + v.y * rhet.y;
      c[k] = v;
      continue;
    }
    const cx<double> r = rp[ig];
    const cx<double> rhog = mk<double>(rhet.x + r.x, rhet.y + r.y);
    const double fpibg = fourpi / (tpiba2 * g2[ig]);
    const cx<double> vcg = mk<double>(fpibg * rhog.x, fpibg * rhog.y);
    eh += 0.5 * (vcg.x * rhog.x + vcg.y * rhog.y);
    e2 += 0.5 * fpibg * (rhet.x * rhet.x + rhet.y * rhet.y);
    g2p += 0.5 * fpibg * (r.x * r.x + r.y * r.y);
    ep += rhet.x * v.x + rhet.y * v.y;
    c[k] = mk<double>(vcg.x + v.x, vcg.y + v.y);
  }
  block_sum(eh, sh, part + 0 * kScfBlocks + blockIdx.x);
  __syncthreads();
  block_sum(ep, sh, part + 1 * kScfBlocks + blockIdx.x);
  __syncthreads();
  block_sum(e2, sh, part + 2 * kScfBlocks + blockIdx.x);
  __syncthreads();
  block_sum(g2p, sh, part + 3 * kScfBlocks + blockIdx.x);
}

// c = (a - b) + 0i  (b may be null)
__global__ void k_real_to_cx(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                             cx<double>* __restrict__ c) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    c[i] = mk<double>(b ? a[i] - b[i] : a[i], 0.0);
}

// KerkerRhoMix.cpp:26-35: on the sphere, c *= beta gg / (gg + q0)
__global__ void k_kerker_filter(cx<double>* __restrict__ c, const int* __restrict__ idx,
                                const double* __restrict__ xkg, int ngw, double beta, double q0) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ngw) return;
  const double gg = xkg[ig];
  const double f = beta * gg / (gg + q0);
  cx<double>& v = c[idx[ig]];
  v = mk<double>(f * v.x, f * v.y);
}

// out = Re(c)
__global__ void k_cx_real(const cx<double>* __restrict__ c, size_t n, double* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = c[i].x;
}

// Linear combination out = sum_k w_k v_k (up to kMaxTerms vectors; out may alias one of
// them: each point is read before it is written), with part = sum (p - q)^2 when p, q set
__global__ __launch_bounds__(kScfThreads) void k_comb(Comb cb, size_t n, double* __restrict__ part) {
  __shared__ double sh[kScfThreads];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kScfThreads + threadIdx.x; i < n; i += (size_t)kScfBlocks * kScfThreads) {
    if (cb.p) {
      const double d = cb.p[i] - (cb.q ? cb.q[i] : 0.0);
      s += d * d;
    }
    double v = 0.0;
    for (int k = 0; k < cb.nt; ++k) v += cb.w[k] * cb.v[k][i];
    cb.out[i] = v;
    if (cb.out2) cb.out2[i] = v;
  }
  if (part) block_sum(s, sh, part + blockIdx.x);
}

// NormalizeRho.cpp:17-21: negatives to zero, partial sums
__global__ __launch_bounds__(kScfThreads) void k_clip_sum(double* __restrict__ rho, size_t n,
                                                          double* __restrict__ part) {
  __shared__ double sh[kScfThreads];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kScfThreads + threadIdx.x; i < n; i += (size_t)kScfBlocks * kScfThreads) {
    double r = rho[i];
    if (r < 0.0) {
      r = 0.0;
      rho[i] = 0.0;
    }
    s += r;
  }
  block_sum(s, sh, part + blockIdx.x);
}

// x = x * num / den, left to right
__global__ void k_scale(double* __restrict__ x, size_t n, double num, double den) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = x[i] * num / den;
}

// the FP32 copy of the local potential (rounded to nearest, as the host conversion)
__global__ void k_to_float(const double* __restrict__ x, size_t n, float* __restrict__ y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

// EvalENL.cpp:36-40: row r of C [rows][ld]: sum_ist focc[ist] |C[r][ist]|^2
__global__ __launch_bounds__(kScfThreads) void k_enl_rows(const cx<double>* __restrict__ C, int ld, int nb,
                                                          const double* __restrict__ focc,
                                                          double* __restrict__ rows) {
  __shared__ double sh[kScfThreads];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < nb; j += kScfThreads) {
    const cx<double> v = C[(size_t)r * ld + j];
    s += focc[j] * (v.x * v.x + v.y * v.y);
  }
  block_sum(s, sh, rows + r);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
dim3 stride_grid(size_t n) {
  const size_t b = (n + 255) / 256;
  return dim3((unsigned)(b < 4096 ? (b ? b : 1) : 4096));
}
}  // namespace

hipError_t xc_pz(const double* rho, size_t n, double* vxc, double* eps, unsigned char* neg,
                 const cx<double>* c, double* out, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_xc, dim3(kScfBlocks), dim3(kScfThreads), 0, s, rho, n, vxc, eps, neg, c, out, part);
  return hipGetLastError();
}
hipError_t xc_pz_fix(const double* rho, size_t n, const unsigned char* neg, double* vxc, double* eps,
                     const cx<double>* c, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_xc_fix, stride_grid(n), dim3(256), 0, s, rho, n, neg, vxc, eps, c, out);
  return hipGetLastError();
}
hipError_t dot_partials(const double* x, const double* y, size_t n, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(kScfBlocks), dim3(kScfThreads), 0, s, x, y, n, part);
  return hipGetLastError();
}
hipError_t energy_g(cx<double>* c, const int* gidx, const double* g2, const cx<double>* vp,
                    const cx<double>* rp, int ng, double tpiba2, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_g, dim3(kScfBlocks), dim3(kScfThreads), 0, s, c, gidx, g2, vp, rp, ng, tpiba2, part);
  return hipGetLastError();
}
hipError_t real_to_cx(const double* a, const double* b, size_t n, cx<double>* c, hipStream_t s) {
  hipLaunchKernelGGL(k_real_to_cx, stride_grid(n), dim3(256), 0, s, a, b, n, c);
  return hipGetLastError();
}
hipError_t kerker_filter(cx<double>* c, const int* idx, const double* xkg, int ngw, double beta, double q0,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_kerker_filter, dim3((ngw + 255) / 256), dim3(256), 0, s, c, idx, xkg, ngw, beta, q0);
  return hipGetLastError();
}
hipError_t cx_real(const cx<double>* c, size_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_cx_real, stride_grid(n), dim3(256), 0, s, c, n, out);
  return hipGetLastError();
}
hipError_t comb(const Comb& cb, size_t n, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_comb, dim3(kScfBlocks), dim3(kScfThreads), 0, s, cb, n, part);
  return hipGetLastError();
}
hipError_t clip_sum(double* rho, size_t n, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_clip_sum, dim3(kScfBlocks), dim3(kScfThreads), 0, s, rho, n, part);
  return hipGetLastError();
}
hipError_t scale_vec(double* x, size_t n, double num, double den, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, stride_grid(n), dim3(256), 0, s, x, n, num, den);
  return hipGetLastError();
}
hipError_t to_float(const double* x, size_t n, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_float, stride_grid(n), dim3(256), 0, s, x, n, y);
  return hipGetLastError();
}
hipError_t enl_rows(const cx<double>* C, int ld, int rows, int nb, const double* focc, double* out,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_enl_rows, dim3(rows), dim3(kScfThreads), 0, s, C, ld, nb, focc, out);
  return hipGetLastError();
}

}  // namespace pwh
