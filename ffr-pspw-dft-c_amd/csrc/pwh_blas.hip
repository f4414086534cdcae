// pwh_blas.hip -- small HBM-bound kernels around the hot path: precision conversion,
// the reference-order <-> sphere-order permutation, the Lanczos reductions
// (lanczos.cpp:28-95 zdotc / zaxpy / zdscal), the subspace fix-ups (symmetrize,
// lobpcg_cpu.cpp:15-27) and the per-band residual norms (ChebySCFLoop.cpp:415-433).
#include "pwh_internal.h"
#include "pwh_blas.h"

namespace pwh {

template <typename Tin, typename Tout>
__global__ void k_convert(const cx<Tin>* __restrict__ in, cx<Tout>* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mk<Tout>((Tout)in[i].x, (Tout)in[i].y);
}

template <typename Tin, typename Tout>
hipError_t convert_cx(const cx<Tin>* in, cx<Tout>* out, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}
template hipError_t convert_cx<double, float>(const cx<double>*, cx<float>*, size_t, hipStream_t);
template hipError_t convert_cx<float, double>(const cx<float>*, cx<double>*, size_t, hipStream_t);

__global__ void k_perm_in(const cx<double>* __restrict__ ref, int ldr, const int* __restrict__ perm,
                          int ngw, int ngw_pad, cx<double>* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= ngw_pad) return;
  out[(size_t)b * ngw_pad + k] = (k < ngw) ? ref[(size_t)b * ldr + perm[k]] : mk<double>(0, 0);
}

__global__ void k_perm_out(const cx<double>* __restrict__ in, int ngw_pad, const int* __restrict__ perm,
                           int ngw, cx<double>* __restrict__ ref, int ldr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= ngw) return;
  ref[(size_t)b * ldr + perm[k]] = in[(size_t)b * ngw_pad + k];
}

hipError_t permute_in_f64(const cx<double>* ref, int ldr, const int* perm, int ngw, int ngw_pad,
                          cx<double>* out, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_perm_in, dim3((ngw_pad + 255) / 256, nb), dim3(256), 0, s, ref, ldr, perm, ngw,
                     ngw_pad, out);
  return hipGetLastError();
}

hipError_t permute_out_f64(const cx<double>* in, int ngw_pad, const int* perm, int ngw,
                           cx<double>* ref, int ldr, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_perm_out, dim3((ngw + 255) / 256, nb), dim3(256), 0, s, in, ngw_pad, perm, ngw,
                     ref, ldr);
  return hipGetLastError();
}

// out[i] = sum_k conj(V[i][k]) f[k], one block per i
__global__ void k_multi_dotc(const cx<double>* __restrict__ V, int ld, const cx<double>* __restrict__ f,
                             int n, cx<double>* __restrict__ out) {
  __shared__ double sr[256], si[256];
  const int i = blockIdx.x, t = threadIdx.x;
  const cx<double>* v = V + (size_t)i * ld;
  double ar = 0, ai = 0;
  for (int k = t; k < n; k += 256) {
    const cx<double> a = v[k], b = f[k];
    ar += a.x * b.x + a.y * b.y;
    ai += a.x * b.y - a.y * b.x;
  }
  sr[t] = ar; si[t] = ai;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) { sr[t] += sr[t + w]; si[t] += si[t + w]; }
    __syncthreads();
  }
  if (t == 0) out[i] = mk<double>(sr[0], si[0]);
}

// f[k] += sum_{i<j} coef[i] V[i][k]
__global__ void k_multi_axpy(const cx<double>* __restrict__ V, int ld, int j,
                             const cx<double>* __restrict__ coef, cx<double>* __restrict__ f, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  cx<double> acc = f[k];
  for (int i = 0; i < j; ++i) acc = acc + coef[i] * V[(size_t)i * ld + k];
  f[k] = acc;
}

// y[k] = a * x[k] + b * z[k]   (z may be null)
__global__ void k_axpby(const cx<double>* __restrict__ x, double a, const cx<double>* __restrict__ z,
                        double b, cx<double>* __restrict__ y, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  cx<double> v = a * x[k];
  if (z) v = v + b * z[k];
  y[k] = v;
}

hipError_t multi_dotc(const cx<double>* V, int ld, int j, const cx<double>* f, int n,
                      cx<double>* out, hipStream_t s) {
  hipLaunchKernelGGL(k_multi_dotc, dim3(j), dim3(256), 0, s, V, ld, f, n, out);
  return hipGetLastError();
}
hipError_t multi_axpy(const cx<double>* V, int ld, int j, const cx<double>* coef, cx<double>* f,
                      int n, hipStream_t s) {
  hipLaunchKernelGGL(k_multi_axpy, dim3((n + 255) / 256), dim3(256), 0, s, V, ld, j, coef, f, n);
  return hipGetLastError();
}
hipError_t axpby(const cx<double>* x, double a, const cx<double>* z, double b, cx<double>* y, int n,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3((n + 255) / 256), dim3(256), 0, s, x, a, z, b, y, n);
  return hipGetLastError();
}

// Column-major n x n: A <- (A + A^H)/2 (lobpcg_cpu.cpp:15-27)
__global__ void k_symmetrize(cx<double>* A, int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i > j || j >= n) return;
  const cx<double> a = A[i + (size_t)j * ld], b = A[j + (size_t)i * ld];
  const cx<double> t = mk<double>(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
  A[i + (size_t)j * ld] = t;
  A[j + (size_t)i * ld] = conjg(t);
}
hipError_t symmetrize(cx<double>* A, int n, int ld, hipStream_t s) {
  hipLaunchKernelGGL(k_symmetrize, dim3((n + 63) / 64, n), dim3(64), 0, s, A, n, ld);
  return hipGetLastError();
}

// Column-major: zero the strict lower triangle
__global__ void k_zero_lower(cx<double>* A, int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || i <= j) return;
  A[i + (size_t)j * ld] = mk<double>(0, 0);
}
hipError_t zero_lower(cx<double>* A, int n, int ld, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_lower, dim3((n + 63) / 64, n), dim3(64), 0, s, A, n, ld);
  return hipGetLastError();
}

// res[b] = || H[b] - lam[b] X[b] ||  (lam null -> || H[b] - X[b] ||)
__global__ void k_resid(const cx<double>* __restrict__ H, const cx<double>* __restrict__ X,
                        const double* __restrict__ lam, int ld, int n, double* __restrict__ res) {
  __shared__ double sh[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const double l = lam ? lam[b] : 1.0;
  double acc = 0;
  for (int k = t; k < n; k += 256) {
    const cx<double> h = H[(size_t)b * ld + k], x = X[(size_t)b * ld + k];
    const double rx = h.x - l * x.x, ry = h.y - l * x.y;
    acc += rx * rx + ry * ry;
  }
  sh[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) res[b] = sqrt(sh[0]);
}
// per band b and slice p: charge[b][p] = sum_k |X[b][k]|^2, ksum[b][p] = sum_k kin[k] |X[b][k]|^2
// over the k of slice p (kBandSlices slices of 256-element chunks per band, fixed-order tree
// reduction); the caller sums the slices in order.
__global__ __launch_bounds__(256) void k_band_sums(const cx<double>* __restrict__ X, int ld, int n,
                                                   const double* __restrict__ kin, double* __restrict__ charge,
                                                   double* __restrict__ ksum) {
  __shared__ double s1[256], s2[256];
  const int b = blockIdx.y, sl = blockIdx.x, t = threadIdx.x;
  const cx<double>* x = X + (size_t)b * ld;
  double a = 0.0, c = 0.0;
  for (int k = sl * 256 + t; k < n; k += 256 * kBandSlices) {
    const double m = x[k].x * x[k].x + x[k].y * x[k].y;
    a += m;
    c += kin[k] * m;
  }
  s1[t] = a;
  s2[t] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) { s1[t] += s1[t + h]; s2[t] += s2[t + h]; }
    __syncthreads();
  }
  if (t == 0) {
    charge[(size_t)b * kBandSlices + sl] = s1[0];
    ksum[(size_t)b * kBandSlices + sl] = s2[0];
  }
}

hipError_t band_sums(const cx<double>* X, int ld, int n, const double* kin, int nb, double* charge,
                     double* ksum, hipStream_t s) {
  hipLaunchKernelGGL(k_band_sums, dim3(kBandSlices, nb), dim3(256), 0, s, X, ld, n, kin, charge, ksum);
  return hipGetLastError();
}

hipError_t resid_norms(const cx<double>* H, const cx<double>* X, const double* lam, int ld, int n,
                       int nb, double* res, hipStream_t s) {
  hipLaunchKernelGGL(k_resid, dim3(nb), dim3(256), 0, s, H, X, lam, ld, n, res);
  return hipGetLastError();
}

}  // namespace pwh
