// pwh_trd.hip -- Householder reduction of a Hermitian matrix to real symmetric tridiagonal
// form for the Rayleigh-Ritz eigensolve above the one-workgroup solver's 128 (C4: 432).
//
// The reduction is the unblocked Householder method in the form of LAPACK's zhetd2 with
// uplo = 'L' (Anderson et al., LAPACK Users' Guide): step k reflects column k below the
// subdiagonal, H_k = I - tau_k v_k v_k^H with v_k(k+1) = 1, chosen as zlarfg chooses it
// (beta = -sign(Re alpha) ||(alpha, x)||, tau = (beta - alpha) / beta, v = x / (alpha - beta)),
// then p = tau B v, w = p - (tau/2)(p^H v) v and B <- B - v w^H - w v^H on the trailing
// block B.  Q = H_0 ... H_{n-2}, A = Q T Q^H; rocSOLVER's zstedc solves T, and the
// back-transformation below applies Q to T's eigenvectors.
//
// GPU form: one launch per step; workgroup g (one wave) owns kTrdCols consecutive columns
// of the full Hermitian matrix (both triangles kept up to date).  A launch for step k
//   * rebuilds w_{k-1} from p_{k-1}, v_{k-1}, tau_{k-1} and the workgroups' partial sums
//     of p_{k-1}^H v_{k-1} (every workgroup does this redundantly: no grid-wide sync);
//   * rebuilds column k with step k-1's update applied and forms the reflector of step k
//     from it (redundantly again); workgroup 0 stores v_k, tau_k, d_k, e_k;
//   * applies step k-1's update to its own columns j > k, then forms their entries of
//     p_k = tau_k B v_k as p_k(j) = tau_k sum_r conj(A(r, j)) v_k(r) (B is Hermitian) and
//     its partial of p_k^H v_k.
// Column k is read by every workgroup and written by none in that launch; p, the partial
// sums and the reflectors are double-buffered by step parity, so launches order all
// cross-workgroup traffic.
#include <cfloat>

#include "pwh_internal.h"

namespace pwh {

constexpr int kTrdCols = 4;          // columns per workgroup
constexpr int kTrdMaxN = 1024;

struct TrdWork {
  cx<double>* A;      // [n][n] column-major working copy (full Hermitian)
  cx<double>* V;      // [n][n] reflector k at V[k n + r], rows r > k
  cx<double>* P;      // [2][n] p of the last step by parity
  cx<double>* PS;     // [2][groups] partial sums of p^H v by parity
  cx<double>* tau;    // [n]
  double* d;          // [n]
  double* e;          // [n]
  int n, groups;
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Element i of a lane's register vector (rows lane + 64 i) at a wave-uniform index:
// selected with compile-time indices (no dynamic register indexing), then broadcast from
// the owning lane.
template <int R>
__device__ __forceinline__ cx<double> row_value(const cx<double> (&a)[R], int r) {
  const int i = r >> 6, src = r & 63;
  cx<double> s = mk<double>(0, 0);
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q == i) s = a[q];
  return mk<double>(__shfl(s.x, src), __shfl(s.y, src));
}

// One Householder step on one wave per workgroup: the vectors of the step live in
// registers (lane l holds rows l + 64 q, q < R, so n <= 64 R), every reduction is a wave
// reduction, and all global loads are issued before the first of them.
template <int R>
__global__ __launch_bounds__(64) void k_trd_step(TrdWork t, int k) {
  const int n = t.n, lane = threadIdx.x, g = blockIdx.x;
  const int prev = (k - 1) & 1, cur = k & 1;
  const int j0 = g * kTrdCols;
  // loads: step k-1's partial sums, v and p; column k; the own columns
  cx<double> ps = mk<double>(0, 0);
  if (k > 0)
    for (int i = lane; i < t.groups; i += 64) ps = ps + t.PS[prev * t.groups + i];
  const cx<double> tau_p = k > 0 ? t.tau[k - 1] : mk<double>(0, 0);
  cx<double> vp[R], wp[R], ck[R], own[kTrdCols][R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    const bool in = r >= k && r < n;
    const int rc = in ? r : k;   // a valid row for the clamped loads
    vp[q] = (k > 0 && r > k && in) ? t.V[(size_t)(k - 1) * n + rc] : mk<double>(r == k ? 1.0 : 0.0, 0.0);
    wp[q] = (k > 0 && in) ? t.P[prev * n + rc] : mk<double>(0, 0);
    ck[q] = in ? t.A[(size_t)k * n + rc] : mk<double>(0, 0);
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) {
      const int j = j0 + c;
      own[c][q] = (j > k && j < n && r > k && r < n) ? t.A[(size_t)j * n + r] : mk<double>(0, 0);
    }
  }
  // w_{k-1} = p_{k-1} - 1/2 tau_{k-1} (p^H v) v_{k-1}
  if (k > 0) {
    ps.x = wave_sum(ps.x);
    ps.y = wave_sum(ps.y);
    const cx<double> al = (-0.5) * (tau_p * ps);
#pragma unroll
    for (int q = 0; q < R; ++q) wp[q] = wp[q] + al * vp[q];
    // column k after step k-1
    const cx<double> vk = mk<double>(1, 0), wk = row_value(wp, k);
#pragma unroll
    for (int q = 0; q < R; ++q) ck[q] = ck[q] - vp[q] * conjg(wk) - wp[q] * conjg(vk);
  }
  // the reflector of step k from x = ck[k+1 .. n-1]
  double xn = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    if (r >= k + 2) xn += ck[q].x * ck[q].x + ck[q].y * ck[q].y;
  }
  const double xnorm2 = wave_sum(xn);
  const cx<double> alpha = row_value(ck, k + 1);
  cx<double> tau = mk<double>(0, 0), scale = mk<double>(1, 0);
  double beta = alpha.x;
  if (xnorm2 > 0.0 || alpha.y != 0.0) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = mk<double>((beta - alpha.x) / beta, -alpha.y / beta);
    const double ar = alpha.x - beta, ai = alpha.y, den = ar * ar + ai * ai;
    scale = mk<double>(ar / den, -ai / den);
  }
  const double dk = row_value(ck, k).x;
  cx<double> vk[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    vk[q] = r == k + 1 ? mk<double>(1, 0) : (r >= k + 2 && r < n) ? scale * ck[q] : mk<double>(0, 0);
  }
  if (g == 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = lane + 64 * q;
      if (r >= k + 2 && r < n) t.V[(size_t)k * n + r] = vk[q];
    }
    if (lane == 0) {
      t.tau[k] = tau;
      t.d[k] = dk;
      t.e[k] = beta;
    }
  }
  // own columns j > k: step k-1's update, then p_k(j) = tau sum_r conj(A(r, j)) v_k(r)
  cx<double> part = mk<double>(0, 0);
#pragma unroll
  for (int c = 0; c < kTrdCols; ++c) {
    const int j = j0 + c;
    if (j <= k || j >= n) continue;   // wave-uniform
    const cx<double> wj = conjg(row_value(wp, j)), vj = conjg(row_value(vp, j));
    cx<double> dot = mk<double>(0, 0);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = lane + 64 * q;
      cx<double> a = own[c][q];
      if (k > 0) a = a - vp[q] * wj - wp[q] * vj;
      if (r > k && r < n) t.A[(size_t)j * n + r] = a;
      dot = dot + conjg(a) * vk[q];
    }
    dot.x = wave_sum(dot.x);
    dot.y = wave_sum(dot.y);
    const cx<double> pj = tau * dot;
    if (lane == 0) t.P[cur * n + j] = pj;
    part = part + conjg(pj) * row_value(vk, j);
  }
  if (lane == 0) t.PS[cur * t.groups + g] = part;
}

// d(n-1) after the last step's update
__global__ void k_trd_last(TrdWork t) {
  if (threadIdx.x != 0) return;
  const int n = t.n, k = n - 2, cur = k & 1;
  cx<double> s = mk<double>(0, 0);
  for (int i = 0; i < t.groups; ++i) s = s + t.PS[cur * t.groups + i];
  const cx<double> tau = t.tau[k];
  const cx<double> al = (-0.5) * (tau * s);
  const cx<double> v = (k + 1 == n - 1) ? mk<double>(1, 0) : t.V[(size_t)k * n + n - 1];
  const cx<double> w = t.P[cur * n + n - 1] + al * v;
  const cx<double> a = t.A[(size_t)(n - 1) * n + n - 1] - v * conjg(w) - w * conjg(v);
  t.d[n - 1] = a.x;
}

// Back-transformation Y <- Q Y with Q = H_0 H_1 ... H_{n-2}: the reflectors from the last to
// the first, y <- y - tau_k v_k (v_k^H y).  Columns of Y are independent: one wave per
// column, the column in registers (lane l holds rows l + 64 i), the dot product a wave
// reduction, so no workgroup barrier is needed.  Reflectors with tau = 0 are skipped.
constexpr int kBtRows = kTrdMaxN / 64;
__global__ __launch_bounds__(256) void k_trd_backtr(const cx<double>* __restrict__ V, const cx<double>* __restrict__ tau,
                                                    cx<double>* __restrict__ Y, int n) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= n) return;   // the whole wave
  cx<double>* y = Y + (size_t)col * n;
  cx<double> yr[kBtRows];
#pragma unroll
  for (int i = 0; i < kBtRows; ++i) {
    const int r = lane + 64 * i;
    yr[i] = r < n ? y[r] : mk<double>(0, 0);
  }
  for (int k = n - 2; k >= 0; --k) {
    const cx<double> tk = tau[k];
    if (tk.x == 0.0 && tk.y == 0.0) continue;
    const cx<double>* vk = V + (size_t)k * n;
    cx<double> vr[kBtRows];
    cx<double> s = mk<double>(0, 0);
#pragma unroll
    for (int i = 0; i < kBtRows; ++i) {
      const int r = lane + 64 * i;
      vr[i] = r == k + 1 ? mk<double>(1, 0) : (r > k + 1 && r < n) ? vk[r] : mk<double>(0, 0);
      s = s + conjg(vr[i]) * yr[i];
    }
    s.x = wave_sum(s.x);
    s.y = wave_sum(s.y);
    const cx<double> f = tk * s;
#pragma unroll
    for (int i = 0; i < kBtRows; ++i) yr[i] = yr[i] - vr[i] * f;
  }
#pragma unroll
  for (int i = 0; i < kBtRows; ++i) {
    const int r = lane + 64 * i;
    if (r < n) y[r] = yr[i];
  }
}

// Eigenvalues of the symmetric tridiagonal (d, e) by multisection on Sturm counts: a
// half-wave per eigenvalue tests 32 points per round (the bracket shrinks 33x), from the
// Gershgorin interval until it is 2 eps |lambda| wide.  The count of eigenvalues below x is
// the number of negative pivots of the LDL^T factorisation of T - x I, q_j = d_j - x -
// e_{j-1}^2 / q_{j-1}, with pivots smaller than pivmin moved to -pivmin.
constexpr int kBisPts = 32;
__global__ __launch_bounds__(256) void k_trd_bisect(const double* __restrict__ d, const double* __restrict__ e, int n,
                                                    double* __restrict__ w) {
  __shared__ double dv[kTrdMaxN], e2[kTrdMaxN], bnd[3];
  const int tid = threadIdx.x;
  for (int j = tid; j < n; j += 256) {
    dv[j] = d[j];
    e2[j] = j + 1 < n ? e[j] * e[j] : 0.0;
  }
  __syncthreads();
  if (tid < 64) {
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
    for (int j = tid; j < n; j += 64) {
      const double el = j > 0 ? sqrt(e2[j - 1]) : 0.0, er = j + 1 < n ? sqrt(e2[j]) : 0.0;
      gl = fmin(gl, dv[j] - el - er);
      gu = fmax(gu, dv[j] + el + er);
      emax2 = fmax(emax2, e2[j]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      gl = fmin(gl, __shfl_xor(gl, m));
      gu = fmax(gu, __shfl_xor(gu, m));
      emax2 = fmax(emax2, __shfl_xor(emax2, m));
    }
    if (tid == 0) {
      const double tn = fmax(fabs(gl), fabs(gu));
      const double pad = 2.0 * DBL_EPSILON * tn * n + DBL_MIN;
      bnd[0] = gl - pad;
      bnd[1] = gu + pad;
      bnd[2] = fmax(emax2, 1.0) * (DBL_MIN / DBL_EPSILON);
    }
  }
  __syncthreads();
  const int i = blockIdx.x * (256 / kBisPts) + (tid >> 5), s = tid & 31;
  if (i >= n) return;
  const double pivmin = bnd[2];
  double lo = bnd[0], hi = bnd[1];
  for (int round = 0; round < 60; ++round) {
    if (!(hi - lo > 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + DBL_MIN)) break;
    const double x = lo + (hi - lo) * (double)(s + 1) / (double)(kBisPts + 1);
    int c = 0;
    double q = dv[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
    for (int j = 1; j < n; ++j) {
      q = dv[j] - x - e2[j - 1] / q;
      if (fabs(q) < pivmin) q = -pivmin;
      if (q < 0.0) ++c;
    }
    double nlo = c <= i ? x : lo, nhi = c > i ? x : hi;
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
      nlo = fmax(nlo, __shfl_xor(nlo, m));
      nhi = fmin(nhi, __shfl_xor(nhi, m));
    }
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  if (s == 0) w[i] = 0.5 * (lo + hi);
}

hipError_t tridiag_eigenvalues(const double* d, const double* e, int n, double* w, hipStream_t s) {
  if (n < 1 || n > kTrdMaxN) return hipErrorInvalidValue;
  const int per = 256 / kBisPts;
  hipLaunchKernelGGL(k_trd_bisect, dim3((n + per - 1) / per), dim3(256), 0, s, d, e, n, w);
  return hipGetLastError();
}

size_t trd_work_elems(int n) {
  const int groups = (n + kTrdCols - 1) / kTrdCols;
  return (size_t)2 * n * n + 2 * (size_t)n + 2 * (size_t)groups;
}

int trd_max_n() { return kTrdMaxN; }

hipError_t tridiag_backtransform(const cx<double>* work, const cx<double>* tau, cx<double>* Y, int n, hipStream_t s) {
  if (n < 3 || n > kTrdMaxN) return hipErrorInvalidValue;
  const cx<double>* V = work + (size_t)n * n;
  hipLaunchKernelGGL(k_trd_backtr, dim3((n + 3) / 4), dim3(256), 0, s, V, tau, Y, n);
  return hipGetLastError();
}

hipError_t tridiag_householder(const cx<double>* A, int n, cx<double>* work, double* d, double* e, cx<double>* tau,
                               hipStream_t s) {
  if (n < 3 || n > kTrdMaxN) return hipErrorInvalidValue;
  TrdWork t;
  t.n = n;
  t.groups = (n + kTrdCols - 1) / kTrdCols;
  t.A = work;
  t.V = t.A + (size_t)n * n;
  t.P = t.V + (size_t)n * n;
  t.PS = t.P + 2 * (size_t)n;
  t.tau = tau;
  t.d = d;
  t.e = e;
  hipError_t err = hipMemcpyAsync(t.A, A, sizeof(cx<double>) * n * n, hipMemcpyDeviceToDevice, s);
  if (err != hipSuccess) return err;
  for (int k = 0; k <= n - 2; ++k) {
    if (n <= 512) hipLaunchKernelGGL(k_trd_step<8>, dim3(t.groups), dim3(64), 0, s, t, k);
    else hipLaunchKernelGGL(k_trd_step<16>, dim3(t.groups), dim3(64), 0, s, t, k);
  }
  hipLaunchKernelGGL(k_trd_last, dim3(1), dim3(64), 0, s, t);
  return hipGetLastError();
}

}  // namespace pwh
