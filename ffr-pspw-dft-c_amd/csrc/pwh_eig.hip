// pwh_eig.hip -- FP64 Hermitian eigensolver for the Rayleigh-Ritz matrices (n <= 128):
// the job of eig_zheevd.cpp:5-34 (zheevd 'V','U') in five launches instead of the hundreds
// of small ones a library divide-and-conquer makes at this size.
//
//   1. k_heev_tridiag (1 workgroup): Householder reduction to real tridiagonal form,
//      A = Q T Q^H, Q = H_0 ... H_{n-2}, H_k = I - tau_k v_k v_k^H (the zhetd2 'L'
//      recurrence: p = tau A v, w = p - tau/2 (p^H v) v, A -= v w^H + w v^H).  The matrix
//      lives in registers: lane (tr, tc) = (tid / 32, tid % 32) holds rows tr + 16a and
//      columns tc + 32b, so the row sums of A v are half-wave shuffles; blocks of the
//      finished leading rows and columns are skipped.
//   2. k_heev_bisect (n / 16 workgroups): eigenvalues of T by multisection on Sturm counts,
//      a half-wave per eigenvalue testing 32 points of its bracket per round (the bracket
//      shrinks 33x per round).
//   3. k_heev_invit (n / 16 workgroups): eigenvectors of T by inverse iteration
//      (T - lambda I = P L U with partial pivoting, three solves), a lane per eigenvalue,
//      factors in LDS.
//   4. k_heev_orth (1 workgroup): every multiplet (a run of eigenvalues with gaps below
//      1e-6 ||T||) orthonormalised member by member (modified Gram-Schmidt, two passes, a
//      half-wave per member), then a Loewdin step Z <- Z - (Z Z^T - I) Z / 2 removes the
//      ~eps ||T|| / gap overlaps left between the other vectors.
//   5. k_heev_backtr (n / 16 workgroups): Y = Q Z, a half-wave per column, Householder
//      vectors staged in LDS.
#include <cfloat>

#include "pwh_internal.h"

namespace pwh {

constexpr int kEigN = 128;                  // largest matrix
constexpr int kEigT = 512;                  // lanes (2 waves per SIMD: up to 256 registers each)
constexpr int kEigG = kEigT / 32;           // half-wave groups (16)
constexpr int kEigRA = kEigN / kEigG;       // row blocks of a lane in phase 1 (8)
constexpr int kEigCB = kEigN / 32;          // column blocks of a lane in phase 1 (4)
constexpr int kBisS = 32;                   // multisection points per eigenvalue per round
constexpr int kInvE = 16;                   // eigenvalues per inverse-iteration workgroup

struct EigScratch {
  double* V;     // [kEigN][kEigN] complex, V[k*kEigN + i] = v_k[i]
  double* Z;     // [kEigN][kEigN] real, Z[i*kEigN + k] = component k of eigenvector i
  double* lu;    // 5 x [kEigN][kEigN]: u0, u1, u2, l, piv of each lane's factorisation
  double* x;     // [kEigN][kEigN] inverse-iteration iterate of each lane
  double* tau;   // [kEigN] complex
  double* d;     // [kEigN] diagonal of T
  double* e;     // [kEigN] off-diagonal of T
  double* lam;   // [kEigN] eigenvalues of the scaled T
  double* st;    // [4] scale applied to T, ||T|| (scaled), Gershgorin low, high
};

__device__ __forceinline__ cx<double> cmul(cx<double> a, cx<double> b) {
  return mk<double>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cx<double> cmulc(cx<double> a, cx<double> b) {   // conj(a) * b
  return mk<double>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ cx<double> shfl_xor_cx(cx<double> v, int m) {
  return mk<double>(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}
// Cross-lane move of a double through DPP (CTRL: quad_perm / row_ror patterns, every
// source lane valid).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// sum over the 32 lanes of a half-wave: quads (xor 1, xor 2), rows of 16 (rotate 4, 8),
// then the two rows (one permute)
__device__ __forceinline__ double halfwave_sum_dpp(double v) {
  v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_f64<0x124>(v);   // row_ror:4
  v += dpp_f64<0x128>(v);   // row_ror:8
  return v + __shfl_xor(v, 16);
}

// value of lane ^ M (M = 1, 2, 4, 8; tc = lane index) within a row of 16 lanes, by DPP:
// quad permutations for 1 and 2, a rotation by 8, and for 4 a rotation by 4 (lanes with
// bit 2 set, whose partner is 4 below) or by 12 (the others)
template <int M>
__device__ __forceinline__ cx<double> xor_row_cx(cx<double> v, int tc) {
  if constexpr (M == 1) {
    return mk<double>(dpp_f64<0xB1>(v.x), dpp_f64<0xB1>(v.y));
  } else if constexpr (M == 2) {
    return mk<double>(dpp_f64<0x4E>(v.x), dpp_f64<0x4E>(v.y));
  } else if constexpr (M == 8) {
    return mk<double>(dpp_f64<0x128>(v.x), dpp_f64<0x128>(v.y));
  } else {
    static_assert(M == 4, "xor distance 1, 2, 4 or 8");
    const bool up = (tc >> 2) & 1;
    const cx<double> a = mk<double>(dpp_f64<0x124>(v.x), dpp_f64<0x124>(v.y));   // row_ror:4
    const cx<double> b = mk<double>(dpp_f64<0x12C>(v.x), dpp_f64<0x12C>(v.y));   // row_ror:12
    return up ? a : b;
  }
}
// Row sums over the 32 lanes of a half-wave for 8 rows at once (reduce-scatter): on return
// the lane holds the complete sum of row 4*b4 + 2*b3 + b2 (b_k = bit k of its index tc in
// the half-wave) -- 4 + 2 + 1 + 2 exchanges instead of 8 x 5; all but the first (across
// the two rows of 16) are DPP moves.
__device__ __forceinline__ cx<double> reduce_scatter8(const cx<double> (&v)[8], int tc, int& row) {
  const bool h4 = (tc >> 4) & 1, h3 = (tc >> 3) & 1, h2 = (tc >> 2) & 1;
  cx<double> a4[4], a2[2];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const cx<double> send = h4 ? v[r] : v[r + 4], keep = h4 ? v[r + 4] : v[r];
    a4[r] = keep + shfl_xor_cx(send, 16);
  }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const cx<double> send = h3 ? a4[r] : a4[r + 2], keep = h3 ? a4[r + 2] : a4[r];
    a2[r] = keep + xor_row_cx<8>(send, tc);
  }
  const cx<double> send = h2 ? a2[0] : a2[1], keep = h2 ? a2[1] : a2[0];
  cx<double> t = keep + xor_row_cx<4>(send, tc);
  t = t + xor_row_cx<2>(t, tc);
  t = t + xor_row_cx<1>(t, tc);
  row = 4 * (int)h4 + 2 * (int)h3 + (int)h2;
  return t;
}

// Workgroup barrier ordering LDS only: global stores still in flight are not waited for
// (the tridiagonal reduction's Householder vectors are read by a later kernel).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// packed offset of v_k[k+1] in the lower triangle of V (v_k has n - 1 - k entries)
__device__ __forceinline__ int vpack_off(int k, int n) { return k * (n - 1) - (k * (k - 1)) / 2; }

// Phase 1.  A: n x n column-major (lda), upper triangle read (Hermitian completion).
__global__ __launch_bounds__(kEigT) void k_heev_tridiag(const cx<double>* __restrict__ A, int lda, int n,
                                                        int* __restrict__ info, EigScratch sc) {
  __shared__ cx<double> vx[2][kEigN];   // x, then v, of the current step (double-buffered)
  __shared__ cx<double> pv[kEigN];      // p = tau A v
  __shared__ cx<double> gpart[kEigG];   // per-row-group partial sums of p^H v
  __shared__ cx<double> tau_s[kEigN];
  __shared__ double nrm2p[kEigG];       // per-row-group partial sums of |x_i|^2, i >= k + 2
  const int tid = threadIdx.x;
  const int tr = tid >> 5, tc = tid & 31, lane = tid & 63;
  if (tid == 0) info[0] = 0;
  // v and p are zero outside the trailing block (indices <= k, and >= n), so the products
  // below need no masks: both v buffers and p start at zero, and each step clears the two
  // v entries and the p entry that fall out of the block
  for (int i = tid; i < kEigN; i += kEigT) {
    vx[0][i] = mk<double>(0.0, 0.0);
    vx[1][i] = mk<double>(0.0, 0.0);
    pv[i] = mk<double>(0.0, 0.0);
  }
  lds_barrier();

  // ---- phase 1: load (upper triangle, Hermitian completion), tridiagonalise ----------
  cx<double> a[kEigRA][kEigCB];
#pragma unroll
  for (int ia = 0; ia < kEigRA; ++ia)
#pragma unroll
    for (int jb = 0; jb < kEigCB; ++jb) {
      const int i = tr + kEigG * ia, j = tc + 32 * jb;
      cx<double> v = mk<double>(0.0, 0.0);
      if (i < n && j < n) {
        if (i < j) v = A[i + (size_t)j * lda];
        else if (i > j) v = conjg(A[j + (size_t)i * lda]);
        else v = mk<double>(A[i + (size_t)i * lda].x, 0.0);
      }
      a[ia][jb] = v;
    }

  for (int k = 0; k + 1 < n; ++k) {
    const int cur = k & 1;
    // x = A[k+1:, k] from the lanes holding column k; d_k from the diagonal's owner
    if (tc == (k & 31)) {
      // column block k >> 5 selected with compile-time indices (no dynamic register indexing);
      // the owners also sum |x_i|^2 below the subdiagonal for the Householder norm
      double part = 0.0;
#pragma unroll
      for (int jb = 0; jb < kEigCB; ++jb) {
        if (jb != (k >> 5)) continue;
#pragma unroll
        for (int ia = 0; ia < kEigRA; ++ia) {
          const int i = tr + kEigG * ia;
          const cx<double> x = a[ia][jb];
          if (i > k && i < n) vx[cur][i] = x;
          if (i > k + 1 && i < n) part = fma(x.x, x.x, fma(x.y, x.y, part));
          if (i == k) sc.d[k] = x.x;
        }
      }
      nrm2p[tr] = part;
    }
    if (tid == 0) {
      // this buffer was last read in step k - 2, before step k - 1's first barrier
      vx[cur][k] = mk<double>(0.0, 0.0);
      if (k > 0) vx[cur][k - 1] = mk<double>(0.0, 0.0);
    }
    lds_barrier();
    // p_k (written last step) is cleared only now: until the barrier above, other waves may
    // still read it in last step's rank-2 update.  Nothing reads p between this barrier and
    // the next, which orders the clear before this step's reads.
    if (tid == 0) pv[k] = mk<double>(0.0, 0.0);
    // Householder vector (zlarfg): H^H x = beta e_1, beta real
    if (tid < 64) {
      const cx<double> alpha = vx[cur][k + 1];
      double xn2 = 0.0;
#pragma unroll
      for (int t = 0; t < kEigG; ++t) xn2 += nrm2p[t];
      cx<double> tau = mk<double>(0.0, 0.0), scale = mk<double>(0.0, 0.0);
      double beta = alpha.x;
      if (xn2 != 0.0 || alpha.y != 0.0) {
        const double r = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = alpha.x >= 0.0 ? -r : r;
        const double rb = 1.0 / beta;
        tau = mk<double>((beta - alpha.x) * rb, -alpha.y * rb);
        const cx<double> dnm = mk<double>(alpha.x - beta, alpha.y);      // 1 / (alpha - beta)
        const double rq = 1.0 / (dnm.x * dnm.x + dnm.y * dnm.y);
        scale = mk<double>(dnm.x * rq, -dnm.y * rq);
      }
      for (int i = k + 2 + lane; i < n; i += 64) {
        const cx<double> v = cmul(vx[cur][i], scale);
        vx[cur][i] = v;
        sc.V[2 * ((size_t)k * kEigN + i)] = v.x;
        sc.V[2 * ((size_t)k * kEigN + i) + 1] = v.y;
      }
      if (lane == 0) {
        vx[cur][k + 1] = mk<double>(1.0, 0.0);
        sc.V[2 * ((size_t)k * kEigN + k + 1)] = 1.0;
        sc.V[2 * ((size_t)k * kEigN + k + 1) + 1] = 0.0;
        tau_s[k] = tau;
        sc.tau[2 * k] = tau.x;
        sc.tau[2 * k + 1] = tau.y;
        sc.e[k] = beta;
      }
    }
    lds_barrier();
    const cx<double> tau = tau_s[k];
    // p_i = tau sum_j A_ij v_j over the trailing block, half-wave reduction over columns
    {
      cx<double> vj[kEigCB];
#pragma unroll
      for (int jb = 0; jb < kEigCB; ++jb) vj[jb] = vx[cur][tc + 32 * jb];   // zero outside the block
      static_assert(kEigRA == 8, "reduce_scatter8 handles 8 row blocks");
      cx<double> rs[kEigRA];
#pragma unroll
      for (int ia = 0; ia < kEigRA; ++ia) {
        cx<double> sum = mk<double>(0.0, 0.0);
        if (kEigG * ia + kEigG - 1 > k) {          // block has rows of the trailing matrix
#pragma unroll
          for (int jb = 0; jb < kEigCB; ++jb)
            if (32 * jb + 31 > k) {   // sum += a v_j as four FMAs
              sum.x = fma(-a[ia][jb].y, vj[jb].y, fma(a[ia][jb].x, vj[jb].x, sum.x));
              sum.y = fma(a[ia][jb].y, vj[jb].x, fma(a[ia][jb].x, vj[jb].y, sum.y));
            }
        }
        rs[ia] = sum;
      }
      int ra;
      const cx<double> rowsum = reduce_scatter8(rs, tc, ra);
      const int i = tr + kEigG * ra;
      cx<double> g = mk<double>(0.0, 0.0);
      if ((tc & 3) == 0 && i > k && i < n) {
        const cx<double> pi = cmul(tau, rowsum);
        pv[i] = pi;
        g = cmulc(pi, vx[cur][i]);
      }
      g = mk<double>(halfwave_sum_dpp(g.x), halfwave_sum_dpp(g.y));
      if (tc == 0) gpart[tr] = g;
    }
    lds_barrier();
    // w = p - tau/2 (p^H v) v ; A -= v w^H + w v^H on the trailing block (v, p from LDS)
    {
      cx<double> gs = mk<double>(0.0, 0.0);
      for (int t = 0; t < kEigG; ++t) gs = gs + gpart[t];
      const cx<double> gam = cmul(mk<double>(-0.5 * tau.x, -0.5 * tau.y), gs);
      cx<double> vj[kEigCB], wj[kEigCB];
#pragma unroll
      for (int jb = 0; jb < kEigCB; ++jb) {   // v and p are zero outside the block, so is w
        const int j = tc + 32 * jb;
        vj[jb] = vx[cur][j];
        wj[jb] = pv[j] + cmul(gam, vj[jb]);
      }
#pragma unroll
      for (int ia = 0; ia < kEigRA; ++ia) {
        if (kEigG * ia + kEigG - 1 <= k) continue;  // finished rows (uniform branch)
        const int i = tr + kEigG * ia;
        const cx<double> vi = vx[cur][i];
        const cx<double> wi = pv[i] + cmul(gam, vi);
#pragma unroll
        for (int jb = 0; jb < kEigCB; ++jb) {
          if (32 * jb + 31 <= k) continue;           // finished columns
          // a -= v_i conj(w_j) + w_i conj(v_j) as eight FMAs; zero outside the trailing block
          const cx<double> wjj = wj[jb], vjj = vj[jb];
          double xr = a[ia][jb].x, xi = a[ia][jb].y;
          xr = fma(-wjj.x, vi.x, xr); xr = fma(-wjj.y, vi.y, xr);
          xr = fma(-vjj.x, wi.x, xr); xr = fma(-vjj.y, wi.y, xr);
          xi = fma(-wjj.x, vi.y, xi); xi = fma(wjj.y, vi.x, xi);
          xi = fma(-vjj.x, wi.y, xi); xi = fma(vjj.y, wi.x, xi);
          a[ia][jb] = mk<double>(xr, xi);
        }
      }
    }
  }
  if (tc == ((n - 1) & 31)) {
#pragma unroll
    for (int jb = 0; jb < kEigCB; ++jb) {
      if (jb != ((n - 1) >> 5)) continue;
#pragma unroll
      for (int ia = 0; ia < kEigRA; ++ia)
        if (tr + kEigG * ia == n - 1) sc.d[n - 1] = a[ia][jb].x;
    }
  }
}

// 1 / q for the Sturm and factorisation recurrences: hardware reciprocal + one Newton step
// (relative error ~1e-16; the recurrences are backward stable, so no exact rounding needed)
__device__ __forceinline__ double frcp(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return r * __builtin_fma(-q, r, 2.0);
}

// number of eigenvalues of T (d, e2 = e^2) below x
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int j = 1; j < n; ++j) {
    q = d[j] - x - e2[j - 1] * frcp(q);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

// T from the tridiagonal reduction into LDS, scaled to unit norm when its norm is far from 1
// (T = 0 gives tnorm = 0); every workgroup computes the same statistics.
__device__ void load_tridiag(int n, const EigScratch& sc, double* dv, double* ev, double* stat) {
  const int tid = threadIdx.x, lane = tid & 63;
  for (int t = tid; t < n; t += blockDim.x) {
    dv[t] = sc.d[t];
    ev[t] = t + 1 < n ? sc.e[t] : 0.0;
  }
  __syncthreads();
  if (tid < 64) {
    double dmax = 0.0, emax = 0.0;
    for (int j = lane; j < n; j += 64) {
      dmax = fmax(dmax, fabs(dv[j]));
      emax = fmax(emax, fabs(ev[j]));
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      dmax = fmax(dmax, __shfl_xor(dmax, m));
      emax = fmax(emax, __shfl_xor(emax, m));
    }
    const double tn0 = dmax + 2.0 * emax;
    const double sc_ = (tn0 > 0.0 && (tn0 < 1e-100 || tn0 > 1e100)) ? 1.0 / tn0 : 1.0;
    if (sc_ != 1.0)
      for (int j = lane; j < n; j += 64) { dv[j] *= sc_; ev[j] *= sc_; }
    if (lane == 0) { stat[0] = sc_; stat[1] = tn0 * sc_; }
  }
  __syncthreads();
}

// Phase 2: eigenvalue i = blockIdx.x * 16 + (tid / 32) by a half-wave; lam (scaled units)
__global__ __launch_bounds__(kEigT) void k_heev_bisect(int n, double* __restrict__ w, EigScratch sc) {
  __shared__ double dv[kEigN], ev[kEigN], e2v[kEigN], stat[4];
  __shared__ double pivmin_s;
  const int tid = threadIdx.x, lane = tid & 63;
  load_tridiag(n, sc, dv, ev, stat);
  if (tid < 64) {
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
    for (int j = lane; j < n; j += 64) {
      const double el = j > 0 ? fabs(ev[j - 1]) : 0.0, er = j + 1 < n ? fabs(ev[j]) : 0.0;
      gl = fmin(gl, dv[j] - el - er);
      gu = fmax(gu, dv[j] + el + er);
      if (j + 1 < n) { emax2 = fmax(emax2, ev[j] * ev[j]); e2v[j] = ev[j] * ev[j]; }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      gl = fmin(gl, __shfl_xor(gl, m));
      gu = fmax(gu, __shfl_xor(gu, m));
      emax2 = fmax(emax2, __shfl_xor(emax2, m));
    }
    if (lane == 0) {
      const double tn = fmax(fabs(gl), fabs(gu));
      const double pad = 2.0 * DBL_EPSILON * tn * n + DBL_MIN;
      stat[2] = gl - pad;
      stat[3] = gu + pad;
      pivmin_s = fmax(emax2, 1.0) * (DBL_MIN / DBL_EPSILON);
      if (blockIdx.x == 0) { sc.st[0] = stat[0]; sc.st[1] = stat[1]; }
    }
  }
  __syncthreads();
  const int i = blockIdx.x * (kEigT / kBisS) + (tid >> 5), s = tid & 31;
  if (i >= n) return;
  const double pivmin = pivmin_s;
  double lo = stat[2], hi = stat[3];
  for (int round = 0; round < 40; ++round) {
    if (!(hi - lo > 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + DBL_MIN)) break;
    const double x = lo + (hi - lo) * (double)(s + 1) / (double)(kBisS + 1);
    const int c = sturm_count(dv, e2v, n, x, pivmin);
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
  if (s == 0) {
    const double l = 0.5 * (lo + hi);
    sc.lam[i] = l;
    w[i] = l / stat[0];   // the same value k_heev_orth stores, available before the vectors
  }
}

// Phase 3: eigenvector of T for eigenvalue i = blockIdx.x * 16 + lane (lanes 0..15), by
// inverse iteration with the factors in LDS ([k][lane], conflict-free).  Z[i][k] (global).
__global__ __launch_bounds__(64) void k_heev_invit(int n, int* __restrict__ info, EigScratch sc) {
  extern __shared__ __align__(16) double ism[];
  double* dv = ism;
  double* ev = dv + kEigN;
  double* u0 = ev + kEigN;
  double* u1 = u0 + kEigN * kInvE;
  double* u2 = u1 + kEigN * kInvE;
  double* ll = u2 + kEigN * kInvE;
  double* pw = ll + kEigN * kInvE;
  double* x = pw + kEigN * kInvE;
  double* stat = x + kEigN * kInvE;
  const int tid = threadIdx.x;
  load_tridiag(n, sc, dv, ev, stat);
  const int l = tid, i = blockIdx.x * kInvE + l;
  if (l >= kInvE || i >= n) return;
  const double tnorm = stat[1];
  if (tnorm == 0.0) {   // T = 0 (A = 0): eigenvectors e_i
    for (int k = 0; k < n; ++k) sc.Z[(size_t)i * kEigN + k] = k == i ? 1.0 : 0.0;
    return;
  }
  const double pivtol = DBL_EPSILON * tnorm;
  const double lmb = sc.lam[i];
  // T - lambda I = P L U, row interchanges where the subdiagonal dominates
  double cd = dv[0] - lmb, cs = n > 1 ? ev[0] : 0.0;
  for (int k = 0; k + 1 < n; ++k) {
    const double sub = ev[k], nd = dv[k + 1] - lmb, ns = k + 2 < n ? ev[k + 1] : 0.0;
    const int o = k * kInvE + l;
    if (fabs(cd) >= fabs(sub)) {
      const double m = cd != 0.0 ? sub * frcp(cd) : 0.0;
      u0[o] = cd; u1[o] = cs; u2[o] = 0.0; ll[o] = m; pw[o] = 0.0;
      cd = nd - m * cs;
      cs = ns;
    } else {
      const double m = cd * frcp(sub);
      u0[o] = sub; u1[o] = nd; u2[o] = ns; ll[o] = m; pw[o] = 1.0;
      const double t = cs - m * nd;
      cs = -m * ns;
      cd = t;
    }
  }
  u0[(n - 1) * kInvE + l] = cd;
  // start vector b_k = cos(k (1 + 0.37 i) + i / 2), then three solves (T - lambda I) x = b / |b|
  double nrm = 0.0;
  for (int k = 0; k < n; ++k) {
    const double b = cos((double)k * (1.0 + 0.37 * i) + 0.5 * i);
    x[k * kInvE + l] = b;
    nrm += b * b;
  }
  double inv = 1.0 / sqrt(nrm);
  for (int it = 0; it < 3; ++it) {
    double yk = x[l] * inv;
    for (int k = 0; k + 1 < n; ++k) {
      const int o = k * kInvE + l;
      double yn = x[o + kInvE] * inv;
      if (pw[o] != 0.0) { const double t = yk; yk = yn; yn = t; }
      x[o] = yk;
      yk = yn - ll[o] * yk;
    }
    x[(n - 1) * kInvE + l] = yk;
    double x1 = 0.0, x2 = 0.0;
    nrm = 0.0;
    for (int k = n - 1; k >= 0; --k) {
      const int o = k * kInvE + l;
      double u = u0[o];
      if (fabs(u) < pivtol) u = u < 0.0 ? -pivtol : pivtol;
      double rhs = x[o];
      if (k + 1 < n) rhs -= u1[o] * x1;
      if (k + 2 < n) rhs -= u2[o] * x2;
      const double xv = rhs * frcp(u);
      x[o] = xv;
      x2 = x1;
      x1 = xv;
      nrm += xv * xv;
    }
    inv = 1.0 / sqrt(nrm);
  }
  for (int k = 0; k < n; ++k) sc.Z[(size_t)i * kEigN + k] = x[k * kInvE + l] * inv;
  if (!(nrm > 0.0 && nrm <= DBL_MAX) || !(lmb == lmb)) atomicOr(info, 1);
}

// Phase 4 (one workgroup): multiplets by modified Gram-Schmidt, then Loewdin steps
// Z <- Z - 1/2 (Z Z^T - I) Z while the largest overlap exceeds 1e-14 (one step when it is
// below 1e-7: the step leaves ~E^2); eigenvalues unscaled into w.
__global__ __launch_bounds__(kEigT) void k_heev_orth(int n, double* __restrict__ w, int* __restrict__ info,
                                                     EigScratch sc) {
  extern __shared__ __align__(16) double zs[];     // [kEigN][kEigN + 1]
  constexpr int ZS = kEigN + 1;
  __shared__ int cstart[kEigN];
  __shared__ int maxpos;
  __shared__ unsigned long long emax_bits;
  const int tid = threadIdx.x;
  if (tid == 0) { maxpos = 0; emax_bits = 0ull; }
  for (int t = tid; t < n * n; t += kEigT) {
    const int i = t / n, k = t - i * n;
    zs[i * ZS + k] = sc.Z[(size_t)i * kEigN + k];
  }
  const double scale = sc.st[0], tnorm = sc.st[1];
  const double ortol = 1e-6 * tnorm;
  __syncthreads();
  if (tid < n) {
    int st = tid;
    while (st > 0 && sc.lam[st] - sc.lam[st - 1] <= ortol) --st;
    cstart[tid] = st;
    atomicMax(&maxpos, tid - st);
    w[tid] = sc.lam[tid] / scale;
  }
  __syncthreads();
  const int hw = tid >> 5, hl = tid & 31;
  const int rounds = tnorm == 0.0 ? 0 : maxpos;    // T = 0: the e_i are orthonormal already
  for (int r = 1; r <= rounds; ++r) {
    for (int i = hw; i < n; i += kEigG) {
      const int st = cstart[i];
      if (i - st != r) continue;
      double xv[kEigCB];
#pragma unroll
      for (int c = 0; c < kEigCB; ++c) {
        const int k = hl + 32 * c;
        xv[c] = k < n ? zs[i * ZS + k] : 0.0;
      }
      for (int pass = 0; pass < 2; ++pass) {
        for (int j = st; j < i; ++j) {
          double zj[kEigCB], dot = 0.0;
#pragma unroll
          for (int c = 0; c < kEigCB; ++c) {
            const int k = hl + 32 * c;
            zj[c] = k < n ? zs[j * ZS + k] : 0.0;
            dot += zj[c] * xv[c];
          }
          dot = halfwave_sum_dpp(dot);
#pragma unroll
          for (int c = 0; c < kEigCB; ++c) xv[c] -= dot * zj[c];
        }
      }
      double nn = 0.0;
#pragma unroll
      for (int c = 0; c < kEigCB; ++c) nn += xv[c] * xv[c];
      nn = halfwave_sum_dpp(nn);
      if (!(nn > 0.0 && nn <= DBL_MAX) && hl == 0) atomicOr(info, 1);
      const double sn = 1.0 / sqrt(nn);
#pragma unroll
      for (int c = 0; c < kEigCB; ++c) {
        const int k = hl + 32 * c;
        if (k < n) zs[i * ZS + k] = xv[c] * sn;
      }
    }
    __syncthreads();
  }
  // Loewdin: E = Z Z^T - I for rows i = ti + 16a, columns j = tj + 32b
  double* E = sc.lu;
  const int ti = tid >> 5, tj = tid & 31;
  for (int pass = 0; pass < 3; ++pass) {
    double acc[kEigRA][kEigCB];
#pragma unroll
    for (int a = 0; a < kEigRA; ++a)
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) acc[a][b] = 0.0;
    for (int k = 0; k < n; ++k) {
      double zi[kEigRA], zj[kEigCB];
#pragma unroll
      for (int a = 0; a < kEigRA; ++a) zi[a] = zs[(ti + kEigG * a) * ZS + k];
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) zj[b] = zs[(tj + 32 * b) * ZS + k];
#pragma unroll
      for (int a = 0; a < kEigRA; ++a)
#pragma unroll
        for (int b = 0; b < kEigCB; ++b) acc[a][b] += zi[a] * zj[b];
    }
    double em = 0.0;
#pragma unroll
    for (int a = 0; a < kEigRA; ++a)
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) {
        const int i = ti + kEigG * a, j = tj + 32 * b;
        if (i < n && j < n) {
          const double e = acc[a][b] - (i == j ? 1.0 : 0.0);
          E[(size_t)i * kEigN + j] = e;
          em = fmax(em, fabs(e));
        }
      }
    // non-negative doubles order like their bit patterns
    atomicMax(&emax_bits, (unsigned long long)__double_as_longlong(em));
    __syncthreads();
    const double emx = __longlong_as_double((long long)emax_bits);
    if (!(emx > 1e-14)) break;
    if (!(emx < 0.1)) { if (tid == 0) atomicOr(info, 1); break; }
    // Z'[i][k] = Z[i][k] - 1/2 sum_j E[i][j] Z[j][k] for rows i = ti + 16a, k = tj + 32b
#pragma unroll
    for (int a = 0; a < kEigRA; ++a)
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) acc[a][b] = 0.0;
    for (int j = 0; j < n; ++j) {
      double ej[kEigRA], zk[kEigCB];
#pragma unroll
      for (int a = 0; a < kEigRA; ++a) {
        const int i = ti + kEigG * a;
        ej[a] = i < n ? E[(size_t)i * kEigN + j] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) zk[b] = zs[j * ZS + tj + 32 * b];
#pragma unroll
      for (int a = 0; a < kEigRA; ++a)
#pragma unroll
        for (int b = 0; b < kEigCB; ++b) acc[a][b] += ej[a] * zk[b];
    }
    __syncthreads();    // every lane has read zs before it is overwritten
#pragma unroll
    for (int a = 0; a < kEigRA; ++a)
#pragma unroll
      for (int b = 0; b < kEigCB; ++b) {
        const int i = ti + kEigG * a, k = tj + 32 * b;
        if (i < n && k < n) zs[i * ZS + k] -= 0.5 * acc[a][b];
      }
    if (tid == 0) emax_bits = 0ull;
    __syncthreads();
    if (emx < 1e-7) break;
  }
  for (int t = tid; t < n * n; t += kEigT) {
    const int i = t / n, k = t - i * n;
    sc.Z[(size_t)i * kEigN + k] = zs[i * ZS + k];
  }
}

// Phase 4: A <- Q Z (the eigenvectors of A, column-major, lda).  Columns are independent:
// a workgroup takes kBtCols of them, one half-wave per column (lane r holds rows r + 32a).
// The Householder vectors are staged in LDS first (packed lower triangle, n (n-1) / 2
// complex), so the chain of n - 1 reflections reads no global memory.
constexpr int kBtCols = kEigT / 32;
__global__ __launch_bounds__(kEigT) void k_heev_backtr(cx<double>* __restrict__ A, int lda, int n,
                                                       EigScratch sc) {
  extern __shared__ __align__(16) unsigned char bsm[];
  cx<double>* vp = reinterpret_cast<cx<double>*>(bsm);
  cx<double>* tp = vp + kEigN * (kEigN - 1) / 2;
  const int tid = threadIdx.x;
  for (int k = 0; k + 1 < n; ++k) {
    const int o = vpack_off(k, n), m = n - 1 - k;
    for (int t = tid; t < m; t += kEigT) {
      const size_t g = 2 * ((size_t)k * kEigN + k + 1 + t);
      vp[o + t] = mk<double>(sc.V[g], sc.V[g + 1]);
    }
  }
  for (int k = tid; k + 1 < n; k += kEigT) tp[k] = mk<double>(sc.tau[2 * k], sc.tau[2 * k + 1]);
  __syncthreads();
  const int r = tid & 31, j = blockIdx.x * kBtCols + (tid >> 5);
  if (j >= n) return;
  cx<double> y[kEigCB];
#pragma unroll
  for (int a = 0; a < kEigCB; ++a) {
    const int i = r + 32 * a;
    y[a] = mk<double>(i < n ? sc.Z[(size_t)j * kEigN + i] : 0.0, 0.0);
  }
  for (int k = n - 2; k >= 0; --k) {
    const cx<double> tk = tp[k];
    const int o = vpack_off(k, n) - (k + 1);
    cx<double> v[kEigCB];
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int a = 0; a < kEigCB; ++a) {
      const int i = r + 32 * a;
      const bool in = i > k && i < n;
      const cx<double> t = vp[in ? o + i : 0];
      v[a] = in ? t : mk<double>(0.0, 0.0);
      // conj(v) y
      sr += v[a].x * y[a].x + v[a].y * y[a].y;
      si += v[a].x * y[a].y - v[a].y * y[a].x;
    }
    sr = halfwave_sum_dpp(sr);
    si = halfwave_sum_dpp(si);
    const cx<double> sv = cmul(tk, mk<double>(sr, si));
#pragma unroll
    for (int a = 0; a < kEigCB; ++a) y[a] = y[a] - cmul(v[a], sv);
  }
#pragma unroll
  for (int a = 0; a < kEigCB; ++a) {
    const int i = r + 32 * a;
    if (i < n) A[i + (size_t)j * lda] = y[a];
  }
}

// V (2 n^2), Z (n^2), lu (n^2), tau (2 n), d, e, lam (n each), st (4, padded to 8)
constexpr size_t kHeevScratchDoubles = (size_t)kEigN * kEigN * (2 + 1 + 1) + 6 * kEigN + 8;
size_t heev_small_scratch_doubles() { return kHeevScratchDoubles; }

// The solver's scratch for callers that cannot allocate (pwh_eig_zheevd_noalloc): a fixed
// 530 KB array in device memory, one copy per GPU, held until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns on each device;
// every call on a GPU shares that GPU's copy.
static __device__ __align__(256) double scratch[kHeevScratchDoubles];

constexpr int kMaxDevices = 64;

hipError_t heev_small_static_scratch(double** ptr) {
  static double* addr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = static_cast<double*>(p);
  }
  *ptr = addr[dev];
  return hipSuccess;
}

static EigScratch eig_scratch(double* scratch) {
  EigScratch sc;
  sc.V = scratch;                               // 2 kEigN^2
  sc.Z = sc.V + 2 * kEigN * kEigN;              // kEigN^2
  sc.lu = sc.Z + kEigN * kEigN;                 // kEigN^2 (the Loewdin E)
  sc.tau = sc.lu + kEigN * kEigN;               // 2 kEigN
  sc.d = sc.tau + 2 * kEigN;
  sc.e = sc.d + kEigN;
  sc.lam = sc.e + kEigN;
  sc.st = sc.lam + kEigN;                       // 4 (+ padding)
  return sc;
}

hipError_t launch_heev_small_values(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                                    hipStream_t s) {
  if (n < 1 || n > kEigN || lda < n) return hipErrorInvalidValue;
  const EigScratch sc = eig_scratch(scratch);
  const int per_bis = kEigT / kBisS;
  hipLaunchKernelGGL(k_heev_tridiag, dim3(1), dim3(kEigT), 0, s, A, lda, n, info, sc);
  hipLaunchKernelGGL(k_heev_bisect, dim3((n + per_bis - 1) / per_bis), dim3(kEigT), 0, s, n, w, sc);
  return hipGetLastError();
}

hipError_t launch_heev_small_vectors(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                                     hipStream_t s) {
  if (n < 1 || n > kEigN || lda < n) return hipErrorInvalidValue;
  const EigScratch sc = eig_scratch(scratch);
  constexpr size_t ism = sizeof(double) * ((size_t)2 * kEigN + 6 * (size_t)kEigN * kInvE + 8);
  constexpr size_t osm = sizeof(double) * (size_t)kEigN * (kEigN + 1);
  constexpr size_t bsm = sizeof(cx<double>) * ((size_t)kEigN * (kEigN - 1) / 2 + kEigN);
  static bool attr = false;
  if (!attr) {
    hipError_t e;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_invit),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)ism)) != hipSuccess ||
        (e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_orth),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)osm)) != hipSuccess ||
        (e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_backtr),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bsm)) != hipSuccess)
      return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_heev_invit, dim3((n + kInvE - 1) / kInvE), dim3(64), ism, s, n, info, sc);
  hipLaunchKernelGGL(k_heev_orth, dim3(1), dim3(kEigT), osm, s, n, w, info, sc);
  hipLaunchKernelGGL(k_heev_backtr, dim3((n + kBtCols - 1) / kBtCols), dim3(kEigT), bsm, s, A, lda, n, sc);
  return hipGetLastError();
}

hipError_t launch_heev_small(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                             hipStream_t s) {
  hipError_t e = launch_heev_small_values(A, lda, n, w, info, scratch, s);
  if (e != hipSuccess) return e;
  return launch_heev_small_vectors(A, lda, n, w, info, scratch, s);
}

}  // namespace pwh
