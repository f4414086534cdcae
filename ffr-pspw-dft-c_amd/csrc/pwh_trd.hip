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
// GPU form: one launch per step; wave j of a step owns column k + 1 + j of the full
// Hermitian matrix (both triangles kept up to date), so only the trailing columns are
// launched.  A launch for step k, in every wave:
//   * rebuilds w_{k-1} from p_{k-1}, v_{k-1}, tau_{k-1} and the waves' partial sums of
//     p_{k-1}^H v_{k-1} (redundantly in every wave: no grid-wide sync);
//   * rebuilds column k with step k-1's update applied and forms the reflector of step k
//     from it (redundantly again); the first wave stores v_k, tau_k, d_k, e_k;
//   * applies step k-1's update to its own column j > k, then forms p_k(j) =
//     tau_k sum_r conj(A(r, j)) v_k(r) (B is Hermitian) and its term of p_k^H v_k.
// Column k is read by every wave and written by none in that launch; p, the partial sums
// and the reflectors are double-buffered by step parity, so launches order all
// cross-workgroup traffic.  Wave sums run as DPP row sums plus four lane reads, and
// values at a wave-uniform row are read with v_readlane (no LDS round trips).
#include <cfloat>
#include <type_traits>

#include "pwh_internal.h"

namespace pwh {

constexpr int kTrdWaves = 4;         // columns (one per wave) per workgroup of a step
constexpr int kTrdMaxN = 1024;

struct TrdWork {
  cx<double>* A;      // [n][n] column-major working copy (full Hermitian)
  cx<double>* V;      // [n][n] reflector k at V[k n + r], rows r > k
  cx<double>* P;      // [2][n] p of the last step by parity
  cx<double>* PS;     // [2][n] terms of p^H v of the last step by parity, by column
  cx<double>* tau;    // [n]
  double* d;          // [n]
  double* e;          // [n]
  int n;
};

// Cross-lane move of a double through DPP (CTRL: quad_perm / row_ror patterns, every
// source lane valid)
template <int CTRL>
__device__ __forceinline__ double trd_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// value of lane l (wave-uniform l)
__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// sum over the wave, the same value in every lane: row sums of 16 by DPP, then the four
// rows' sums read from lanes 0, 16, 32 and 48
__device__ __forceinline__ double wave_sum(double x) {
  x += trd_dpp<0xB1>(x);    // quad_perm [1,0,3,2]
  x += trd_dpp<0x4E>(x);    // quad_perm [2,3,0,1]
  x += trd_dpp<0x124>(x);   // row_ror:4
  x += trd_dpp<0x128>(x);   // row_ror:8
  return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// Element i of a lane's register vector (rows lane + 64 i) at a wave-uniform index:
// selected with compile-time indices (no dynamic register indexing), then read from the
// owning lane.  Called with every lane of the wave active (the selection runs per lane).
template <int R>
__device__ __forceinline__ cx<double> row_value(const cx<double> (&a)[R], int r) {
  const int i = r >> 6, src = r & 63;
  cx<double> s = mk<double>(0, 0);
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q == i) s = a[q];
  return mk<double>(read_lane(s.x, src), read_lane(s.y, src));
}

// One Householder step, one wave per trailing column: the vectors of the step live in
// registers (lane l holds rows l + 64 q, q < R, so n <= 64 R), every reduction is a wave
// reduction, and all global loads are issued before the first of them.
template <int R>
__global__ __launch_bounds__(64 * kTrdWaves) void k_trd_step(TrdWork t, int k) {
  const int n = t.n, lane = threadIdx.x & 63;
  const int j = k + 1 + blockIdx.x * kTrdWaves + (threadIdx.x >> 6);   // the wave's column
  if (j >= n) return;   // the whole wave; the first wave (j = k + 1 < n) always stays
  const bool lead = j == k + 1;
  const int prev = (k - 1) & 1, cur = k & 1;
  // loads: step k-1's partial sums (columns k .. n-1), v and p; column k; the own column
  cx<double> psq[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = k + lane + 64 * q;
    psq[q] = (k > 0 && i < n) ? t.PS[prev * n + i] : mk<double>(0, 0);
  }
  const cx<double> tau_p = k > 0 ? t.tau[k - 1] : mk<double>(0, 0);
  cx<double> vp[R], wp[R], ck[R], own[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    const bool in = r >= k && r < n;
    const int rc = in ? r : k;   // a valid row for the clamped loads
    vp[q] = (k > 0 && r > k && in) ? t.V[(size_t)(k - 1) * n + rc] : mk<double>(r == k ? 1.0 : 0.0, 0.0);
    wp[q] = (k > 0 && in) ? t.P[prev * n + rc] : mk<double>(0, 0);
    ck[q] = in ? t.A[(size_t)k * n + rc] : mk<double>(0, 0);
    own[q] = (r > k && r < n) ? t.A[(size_t)j * n + r] : mk<double>(0, 0);
  }
  // w_{k-1} = p_{k-1} - 1/2 tau_{k-1} (p^H v) v_{k-1}
  if (k > 0) {
    cx<double> ps = psq[0];
#pragma unroll
    for (int q = 1; q < R; ++q) ps = ps + psq[q];
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
  cx<double> vk[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    vk[q] = r == k + 1 ? mk<double>(1, 0) : (r >= k + 2 && r < n) ? scale * ck[q] : mk<double>(0, 0);
  }
  if (lead) {
    const double dk = row_value(ck, k).x;
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
  // own column j > k: step k-1's update, then p_k(j) = tau sum_r conj(A(r, j)) v_k(r)
  const cx<double> wj = conjg(row_value(wp, j)), vj = conjg(row_value(vp, j));
  cx<double> dot = mk<double>(0, 0);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    cx<double> a = own[q];
    if (k > 0) a = a - vp[q] * wj - wp[q] * vj;
    if (r > k && r < n) t.A[(size_t)j * n + r] = a;
    dot = dot + conjg(a) * vk[q];
  }
  dot.x = wave_sum(dot.x);
  dot.y = wave_sum(dot.y);
  const cx<double> pj = tau * dot;
  const cx<double> vkj = row_value(vk, j);   // (lane reads: every lane active)
  if (lane == 0) {
    t.P[cur * n + j] = pj;
    t.PS[cur * n + j] = conjg(pj) * vkj;
  }
}

// d(n-1) after the last step's update
__global__ void k_trd_last(TrdWork t) {
  if (threadIdx.x != 0) return;
  const int n = t.n, k = n - 2, cur = k & 1;
  cx<double> s = mk<double>(0, 0);
  for (int i = k + 1; i < n; ++i) s = s + t.PS[cur * n + i];
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
// reduction.  The workgroup's 4 waves share the reflectors: blocks of kb of them (zero
// above row k + 1, one at k + 1) are staged in LDS, the next block's loads in flight while
// the waves apply the current one.  Reflectors with tau = 0 are skipped.
constexpr int kBtWaves = 4;
constexpr int kBtPf = 16;   // staged values per lane: kb n <= kBtPf 64 kBtWaves
static int backtr_kb(int n) { return n <= 512 ? 8 : 4; }
template <int R>
__global__ __launch_bounds__(64 * kBtWaves) void k_trd_backtr(const cx<double>* __restrict__ V,
                                                              const cx<double>* __restrict__ tau,
                                                              cx<double>* __restrict__ Y, int n, int kb) {
  extern __shared__ __align__(16) unsigned char trd_sm[];
  cx<double>* sV = reinterpret_cast<cx<double>*>(trd_sm);   // [kb][n]
  __shared__ cx<double> sT[8];
  const int tid = threadIdx.x, lane = tid & 63;
  const int col = blockIdx.x * kBtWaves + (tid >> 6);
  const bool act = col < n;   // waves past the last column still stage
  cx<double>* y = Y + (size_t)(act ? col : 0) * n;
  cx<double> yr[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int r = lane + 64 * i;
    yr[i] = (act && r < n) ? y[r] : mk<double>(0, 0);
  }
  const int per = kb * n;
  cx<double> pf[kBtPf];
  // block of reflectors k = hi, hi - 1, ..., hi - kb + 1 (slot kk = hi - k)
  auto fetch = [&](int hi) {
#pragma unroll
    for (int q = 0; q < kBtPf; ++q) {
      const int e = tid + 64 * kBtWaves * q;
      cx<double> v = mk<double>(0, 0);
      if (e < per) {
        const int kk = e / n, r = e - kk * n, k = hi - kk;
        if (k >= 0) v = r == k + 1 ? mk<double>(1, 0) : r > k + 1 ? V[(size_t)k * n + r] : mk<double>(0, 0);
      }
      pf[q] = v;
    }
  };
  fetch(n - 2);
  for (int hi = n - 2; hi >= 0; hi -= kb) {
#pragma unroll
    for (int q = 0; q < kBtPf; ++q) {
      const int e = tid + 64 * kBtWaves * q;
      if (e < per) sV[e] = pf[q];
    }
    if (tid < kb) sT[tid] = hi - tid >= 0 ? tau[hi - tid] : mk<double>(0, 0);
    __syncthreads();
    if (hi - kb >= 0) fetch(hi - kb);
    if (act) {
      for (int kk = 0; kk < kb && hi - kk >= 0; ++kk) {
        const cx<double> tk = sT[kk];
        if (tk.x == 0.0 && tk.y == 0.0) continue;
        cx<double> vr[R];
        cx<double> s = mk<double>(0, 0);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int r = lane + 64 * i;
          vr[i] = r < n ? sV[kk * n + r] : mk<double>(0, 0);
          s = s + conjg(vr[i]) * yr[i];
        }
        s.x = wave_sum(s.x);
        s.y = wave_sum(s.y);
        const cx<double> f = tk * s;
#pragma unroll
        for (int i = 0; i < R; ++i) yr[i] = yr[i] - vr[i] * f;
      }
    }
    __syncthreads();
  }
  if (act) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int r = lane + 64 * i;
      if (r < n) y[r] = yr[i];
    }
  }
}

// the row-register count R of a kernel for n: ceil(n / 64) rounded up to an instantiated size
template <typename F>
static void with_rows(int n, F&& f) {
  const int rows = (n + 63) / 64;
  if (rows <= 3) f(std::integral_constant<int, 3>{});
  else if (rows <= 4) f(std::integral_constant<int, 4>{});
  else if (rows <= 5) f(std::integral_constant<int, 5>{});
  else if (rows <= 6) f(std::integral_constant<int, 6>{});
  else if (rows <= 7) f(std::integral_constant<int, 7>{});
  else if (rows <= 8) f(std::integral_constant<int, 8>{});
  else if (rows <= 10) f(std::integral_constant<int, 10>{});
  else if (rows <= 12) f(std::integral_constant<int, 12>{});
  else f(std::integral_constant<int, 16>{});
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

size_t trd_work_elems(int n) { return (size_t)2 * n * n + 4 * (size_t)n; }

int trd_max_n() { return kTrdMaxN; }

hipError_t tridiag_backtransform(const cx<double>* work, const cx<double>* tau, cx<double>* Y, int n, hipStream_t s) {
  if (n < 3 || n > kTrdMaxN) return hipErrorInvalidValue;
  const cx<double>* V = work + (size_t)n * n;
  const int kb = backtr_kb(n);
  const size_t sm = sizeof(cx<double>) * (size_t)kb * n;   // <= 64 KB
  with_rows(n, [&](auto rc) {
    hipLaunchKernelGGL(k_trd_backtr<decltype(rc)::value>, dim3((n + kBtWaves - 1) / kBtWaves), dim3(64 * kBtWaves),
                       sm, s, V, tau, Y, n, kb);
  });
  return hipGetLastError();
}

hipError_t tridiag_householder(const cx<double>* A, int n, cx<double>* work, double* d, double* e, cx<double>* tau,
                               hipStream_t s) {
  if (n < 3 || n > kTrdMaxN) return hipErrorInvalidValue;
  TrdWork t;
  t.n = n;
  t.A = work;
  t.V = t.A + (size_t)n * n;
  t.P = t.V + (size_t)n * n;
  t.PS = t.P + 2 * (size_t)n;
  t.tau = tau;
  t.d = d;
  t.e = e;
  hipError_t err = hipMemcpyAsync(t.A, A, sizeof(cx<double>) * n * n, hipMemcpyDeviceToDevice, s);
  if (err != hipSuccess) return err;
  with_rows(n, [&](auto rc) {
    for (int k = 0; k <= n - 2; ++k) {
      const int blocks = (n - k - 1 + kTrdWaves - 1) / kTrdWaves;
      hipLaunchKernelGGL(k_trd_step<decltype(rc)::value>, dim3(blocks), dim3(64 * kTrdWaves), 0, s, t, k);
    }
  });
  hipLaunchKernelGGL(k_trd_last, dim3(1), dim3(64), 0, s, t);
  return hipGetLastError();
}

}  // namespace pwh
