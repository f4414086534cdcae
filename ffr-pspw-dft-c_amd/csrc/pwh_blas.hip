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

// Grouped conversion (or copy, Tin = Tout) of one array per system: block L belongs to the
// array whose first block is the largest one <= L, 256 elements per block.
template <typename Tin, typename Tout>
__global__ void k_convert_group(const GroupConv* __restrict__ g, int ng) {
  const int L = blockIdx.x;
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].blk0 <= L) lo = mid;
    else hi = mid - 1;
  }
  const GroupConv& p = g[lo];
  const size_t i = (size_t)(L - p.blk0) * 256 + threadIdx.x;
  if (i < p.n) {
    const cx<Tin> v = static_cast<const cx<Tin>*>(p.in)[i];
    static_cast<cx<Tout>*>(p.out)[i] = mk<Tout>((Tout)v.x, (Tout)v.y);
  }
}

template <typename Tin, typename Tout>
hipError_t launch_group_convert(const GroupConv* g, int n, int blocks, hipStream_t s) {
  if (n < 1 || blocks < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_convert_group<Tin, Tout>), dim3(blocks), dim3(256), 0, s, g, n);
  return hipGetLastError();
}
template hipError_t launch_group_convert<double, float>(const GroupConv*, int, int, hipStream_t);
template hipError_t launch_group_convert<float, double>(const GroupConv*, int, int, hipStream_t);
template hipError_t launch_group_convert<double, double>(const GroupConv*, int, int, hipStream_t);

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

// f[k] = src[k] - sum_{i<j} c[i] V[i][k], the coefficients read on the device (src may be f)
__global__ void k_multi_axpy_sub(const cx<double>* __restrict__ V, int ld, int j, const cx<double>* __restrict__ c,
                                 const cx<double>* src, cx<double>* f, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  cx<double> acc = src[k];
  for (int i = 0; i < j; ++i) acc = acc - c[i] * V[(size_t)i * ld + k];
  f[k] = acc;
}

// y = x / sqrt(Re nrm[0]) (the reciprocal formed once, as zdscal's beta); thread 0 also
// writes sqrt(Re nrm[0]) to *tpos when tpos is given
__global__ void k_scale_rnorm(const cx<double>* __restrict__ x, const cx<double>* __restrict__ nrm,
                              cx<double>* __restrict__ y, int n, double* __restrict__ tpos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const double r = sqrt(nrm[0].x);
  if (tpos && k == 0) *tpos = r;
  if (k >= n) return;
  y[k] = (1.0 / r) * x[k];
}

// tcol[i] = Re(h[i] + s[i]) for i < j (s null: Re h[i])
__global__ void k_tcol(const cx<double>* __restrict__ h, const cx<double>* __restrict__ s, int j,
                       double* __restrict__ tcol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= j) return;
  tcol[i] = s ? h[i].x + s[i].x : h[i].x;
}

// out[i] = sum_k conj(V[i][k]) f[k] over P slices of k: grid (P, j) writes part[i][p],
// then one thread per i adds its P partials in order (deterministic)
__global__ __launch_bounds__(256) void k_dotc_part(const cx<double>* __restrict__ V, int ld,
                                                   const cx<double>* __restrict__ f, int n, int P,
                                                   cx<double>* __restrict__ part) {
  __shared__ double sr[256], si[256];
  const int p = blockIdx.x, i = blockIdx.y, t = threadIdx.x;
  const int chunk = (n + P - 1) / P, k0 = p * chunk, k1 = min(n, k0 + chunk);
  const cx<double>* v = V + (size_t)i * ld;
  double ar = 0, ai = 0;
  for (int k = k0 + t; k < k1; k += 256) {
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
  if (t == 0) part[(size_t)i * P + p] = mk<double>(sr[0], si[0]);
}
__global__ void k_dotc_fin(const cx<double>* __restrict__ part, int P, int j, cx<double>* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= j) return;
  cx<double> s = mk<double>(0, 0);
  for (int p = 0; p < P; ++p) s = s + part[(size_t)i * P + p];
  out[i] = s;
}
hipError_t multi_dotc_split(const cx<double>* V, int ld, int j, const cx<double>* f, int n, cx<double>* part,
                            int P, cx<double>* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dotc_part, dim3(P, j), dim3(256), 0, s, V, ld, f, n, P, part);
  hipLaunchKernelGGL(k_dotc_fin, dim3((j + 63) / 64), dim3(64), 0, s, part, P, j, out);
  return hipGetLastError();
}

__global__ void k_fill_zero(double* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
hipError_t fill_zero(double* p, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t multi_axpy_sub(const cx<double>* V, int ld, int j, const cx<double>* c, const cx<double>* src,
                          cx<double>* f, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_multi_axpy_sub, dim3((n + 255) / 256), dim3(256), 0, s, V, ld, j, c, src, f, n);
  return hipGetLastError();
}
hipError_t scale_rnorm(const cx<double>* x, const cx<double>* nrm, cx<double>* y, int n, double* tpos,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_scale_rnorm, dim3((n + 255) / 256), dim3(256), 0, s, x, nrm, y, n, tpos);
  return hipGetLastError();
}
hipError_t tcol(const cx<double>* h, const cx<double>* sv, int j, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tcol, dim3((j + 255) / 256), dim3(256), 0, s, h, sv, j, out);
  return hipGetLastError();
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

// Column-major: fill the strict lower triangle with the conjugate of the upper one, and
// keep only the real part of the diagonal (as LAPACK's Hermitian routines read it)
__global__ void k_herm_fill_lower(cx<double>* A, int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || i < j) return;
  if (i == j) A[i + (size_t)i * ld].y = 0.0;
  else A[i + (size_t)j * ld] = conjg(A[j + (size_t)i * ld]);
}
hipError_t herm_fill_lower(cx<double>* A, int n, int ld, hipStream_t s) {
  hipLaunchKernelGGL(k_herm_fill_lower, dim3((n + 63) / 64, n), dim3(64), 0, s, A, n, ld);
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

// Column-major n x n, ld = n: G <- G - I in place, maxabs[0] = the largest |e_ij| of the
// result (infinite when an entry is not a number).  One workgroup per column; the column
// maxima meet in an atomic max on the bit patterns (non-negative doubles order as their
// bits, and infinity above every finite value), so the result does not depend on the order.
__global__ void k_zero_max(double* maxabs) { maxabs[0] = 0.0; }

__global__ __launch_bounds__(256) void k_eye_residual(cx<double>* G, int n, double* maxabs) {
  __shared__ double red[256];
  const int tid = threadIdx.x, j = blockIdx.x;
  double m = 0.0;
  for (int i = tid; i < n; i += 256) {
    cx<double> v = G[i + (size_t)j * n];
    if (i == j) {
      v.x -= 1.0;
      G[i + (size_t)j * n] = v;
    }
    // |e_ij| (the modulus, not the larger component), so n max|e_ij| bounds ||E||_2
    const double a = hypot(v.x, v.y);
    m = (a == a) ? fmax(m, a) : INFINITY;
  }
  red[tid] = m;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (tid < h) red[tid] = fmax(red[tid], red[tid + h]);
    __syncthreads();
  }
  if (tid == 0)
    atomicMax(reinterpret_cast<unsigned long long*>(maxabs), (unsigned long long)__double_as_longlong(red[0]));
}
hipError_t eye_residual(cx<double>* G, int n, double* maxabs, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_max, dim3(1), dim3(1), 0, s, maxabs);
  hipLaunchKernelGGL(k_eye_residual, dim3(n), dim3(256), 0, s, G, n, maxabs);
  return hipGetLastError();
}

// out = a in + b I (column-major n x n, ld = n; out may be in)
__global__ void k_axpy_eye(cx<double>* out, const cx<double>* in, double a, double b, int n) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n) return;
  cx<double> v = in[e];
  v.x *= a;
  v.y *= a;
  if (e % (n + 1) == 0) v.x += b;
  out[e] = v;
}
hipError_t axpy_eye(cx<double>* out, const cx<double>* in, double a, double b, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_eye, dim3((n * n + 255) / 256), dim3(256), 0, s, out, in, a, b, n);
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

// ---------------------------------------------------------------------------
// Block-solver helpers (davidson_cpu.cpp, lobpcg_cpu.cpp).  Per-band reductions run over
// kBandSlices slices of 256-element chunks and are finished in slice order
// (deterministic).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slice_reduce(double v, double* sh, double* out) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

// mode 0: R = lam[b] X - Y (davidson_cpu.cpp:64-70); mode 1: R = Y - X (W = HX - X T,
// lobpcg_cpu.cpp:98-104, with X holding X T).  part[b][slice] = sum |R|^2 over the slice.
__global__ __launch_bounds__(256) void k_resid_block(const cx<double>* __restrict__ X,
                                                     const cx<double>* __restrict__ Y,
                                                     const double* __restrict__ lam, int ld, int n, int mode,
                                                     cx<double>* __restrict__ R, double* __restrict__ part) {
  __shared__ double sh[256];
  const int b = blockIdx.y, sl = blockIdx.x;
  const size_t o = (size_t)b * ld;
  const double l = mode == 0 ? lam[b] : 0.0;
  double acc = 0.0;
  for (int k = sl * 256 + threadIdx.x; k < n; k += 256 * kBandSlices) {
    const cx<double> x = X[o + k], y = Y[o + k];
    const cx<double> r = mode == 0 ? mk<double>(l * x.x - y.x, l * x.y - y.y) : mk<double>(y.x - x.x, y.y - x.y);
    R[o + k] = r;
    acc += r.x * r.x + r.y * r.y;
  }
  slice_reduce(acc, sh, part + (size_t)b * kBandSlices + sl);
}

// part[b][slice] = Re sum conj(X) Y over the slice (zdotc, davidson_cpu.cpp:58-61)
__global__ __launch_bounds__(256) void k_dotc_re(const cx<double>* __restrict__ X, const cx<double>* __restrict__ Y,
                                                 int ld, int n, double* __restrict__ part) {
  __shared__ double sh[256];
  const int b = blockIdx.y, sl = blockIdx.x;
  const size_t o = (size_t)b * ld;
  double acc = 0.0;
  for (int k = sl * 256 + threadIdx.x; k < n; k += 256 * kBandSlices) {
    const cx<double> x = X[o + k], y = Y[o + k];
    acc += x.x * y.x + x.y * y.y;
  }
  slice_reduce(acc, sh, part + (size_t)b * kBandSlices + sl);
}

// out[b] = sum of the slices (sqrt when root)
__global__ void k_finish_slices(const double* __restrict__ part, int nb, int root, double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0.0;
  for (int p = 0; p < kBandSlices; ++p) s += part[(size_t)b * kBandSlices + p];
  out[b] = root ? sqrt(s) : s;
}

hipError_t resid_block(const cx<double>* X, const cx<double>* Y, const double* lam, int ld, int n, int nb,
                       int mode, cx<double>* R, double* part, double* norms, hipStream_t s) {
  hipLaunchKernelGGL(k_resid_block, dim3(kBandSlices, nb), dim3(256), 0, s, X, Y, lam, ld, n, mode, R, part);
  hipLaunchKernelGGL(k_finish_slices, dim3((nb + 63) / 64), dim3(64), 0, s, part, nb, 1, norms);
  return hipGetLastError();
}

hipError_t dotc_re(const cx<double>* X, const cx<double>* Y, int ld, int n, int nb, double* part, double* out,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_dotc_re, dim3(kBandSlices, nb), dim3(256), 0, s, X, Y, ld, n, part);
  hipLaunchKernelGGL(k_finish_slices, dim3((nb + 63) / 64), dim3(64), 0, s, part, nb, 0, out);
  return hipGetLastError();
}

// R[b][k] = (R[b][k] * c_b) * prec[k] with c_b = 1/res[b] when res[b] > machine zero, else 1
// (davidson_cpu.cpp:92-107); res null: c_b = 1 (lobpcg_cpu.cpp:126-131)
__global__ void k_scale_prec(cx<double>* __restrict__ R, int ld, int n, const double* __restrict__ res,
                             const double* __restrict__ prec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= n) return;
  cx<double> v = R[(size_t)b * ld + k];
  if (res) {
    const double r = res[b];
    const double c = r > 2.220446049250313e-16 ? 1.0 / r : 1.0;
    v = mk<double>(v.x * c, v.y * c);
  }
  const double p = prec[k];
  R[(size_t)b * ld + k] = mk<double>(v.x * p, v.y * p);
}
hipError_t scale_prec(cx<double>* R, int ld, int n, int nb, const double* res, const double* prec,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_scale_prec, dim3((n + 255) / 256, nb), dim3(256), 0, s, R, ld, n, res, prec);
  return hipGetLastError();
}

// Column-major M (ld): the leading n x n block becomes diag(vals) (vals null: identity)
__global__ void k_set_diag_block(cx<double>* M, int ld, int n, const double* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  M[i + (size_t)j * ld] = i == j ? mk<double>(vals ? vals[i] : 1.0, 0.0) : mk<double>(0.0, 0.0);
}
hipError_t set_diag_block(cx<double>* M, int ld, int n, const double* vals, hipStream_t s) {
  hipLaunchKernelGGL(k_set_diag_block, dim3((n + 63) / 64, n), dim3(64), 0, s, M, ld, n, vals);
  return hipGetLastError();
}

// Column-major n x n: the strict lower triangle from the upper one, the diagonal real
// (what zheevd with uplo = 'U' reads of a matrix)
__global__ void k_fill_lower(cx<double>* A, int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i > j || j >= n) return;
  if (i == j) {
    A[i + (size_t)i * ld].y = 0.0;
  } else {
    A[j + (size_t)i * ld] = conjg(A[i + (size_t)j * ld]);
  }
}
hipError_t fill_lower(cx<double>* A, int n, int ld, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_lower, dim3((n + 63) / 64, n), dim3(64), 0, s, A, n, ld);
  return hipGetLastError();
}

// eig_zhegv.cpp:40-46: column i of V scaled by 1/sqrt(b[i]) when b[i] > SMALL
__global__ void k_zhegv_scale(cx<double>* V, int n, int ld, const double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const double e = b[j];
  if (e > 12.220446049250313e-16) {
    const double c = 1.0 / sqrt(e);
    cx<double>& v = V[i + (size_t)j * ld];
    v = mk<double>(v.x * c, v.y * c);
  }
}
hipError_t zhegv_scale(cx<double>* V, int n, int ld, const double* b, hipStream_t s) {
  hipLaunchKernelGGL(k_zhegv_scale, dim3((n + 63) / 64, n), dim3(64), 0, s, V, n, ld, b);
  return hipGetLastError();
}

hipError_t resid_norms(const cx<double>* H, const cx<double>* X, const double* lam, int ld, int n,
                       int nb, double* res, hipStream_t s) {
  hipLaunchKernelGGL(k_resid, dim3(nb), dim3(256), 0, s, H, X, lam, ld, n, res);
  return hipGetLastError();
}

}  // namespace pwh
