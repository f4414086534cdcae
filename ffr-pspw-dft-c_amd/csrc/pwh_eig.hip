// pwh_eig.hip -- FP64 Hermitian eigensolver for the Rayleigh-Ritz matrices (n <= 128) in
// one workgroup: the job of eig_zheevd.cpp:5-34 (zheevd 'V','U') without the hundreds
// of small launches a library divide-and-conquer makes at this size.
//
// Three kernels of one 512-lane workgroup each (phases 1, 2-3, 4):
//   1. Householder reduction to real tridiagonal form, A = Q T Q^H with
//      Q = H_0 H_1 ... H_{n-2}, H_k = I - tau_k v_k v_k^H (the zhetd2 'L' recurrence:
//      p = tau A v, w = p - tau/2 (p^H v) v, A -= v w^H + w v^H).  The matrix lives in
//      registers: lane (tr, tc) = (tid / 32, tid % 32) holds rows tr + 16a and columns
//      tc + 32b (32 entries), so the row sums of A v are half-wave shuffles.
//   2. Eigenvalues of T by multisection on Sturm counts: 4 lanes per eigenvalue test 4
//      points of its bracket per round (the bracket shrinks 5x per round).
//   3. Eigenvectors of T by inverse iteration (T - lambda I = P L U with partial
//      pivoting, three solves), all eigenvalues at once; then every cluster (a run of
//      eigenvalues with gaps below 1e-6 ||T||) is orthonormalised member by member
//      (modified Gram-Schmidt, two passes, a half-wave per member), and a Loewdin step
//      Z <- Z - (Z Z^T - I) Z / 2 removes the ~eps ||T|| / gap overlaps that remain.
//   4. Back-transformation Y = Q Z, registers again: lane (r, c) = (tid % 32, tid / 32)
//      holds rows r + 32a and columns c + 16b, so a half-wave owns whole columns and
//      needs no barrier.
#include <cfloat>

#include "pwh_internal.h"

namespace pwh {

constexpr int kEigN = 128;                  // largest matrix
constexpr int kEigT = 512;                  // lanes (2 waves per SIMD: up to 256 registers each)
constexpr int kEigG = kEigT / 32;           // half-wave groups (16)
constexpr int kEigRA = kEigN / kEigG;       // row blocks of a lane in phase 1 (8)
constexpr int kEigCB = kEigN / 32;          // column blocks of a lane in phase 1 (4)
constexpr int kEigSamples = kEigT / kEigN;  // multisection points per eigenvalue per round (4)

struct EigScratch {
  double* V;     // [kEigN][kEigN] complex, V[k*kEigN + i] = v_k[i]
  double* Z;     // [kEigN][kEigN] real, Z[i*kEigN + k] = component k of eigenvector i
  double* lu;    // 5 x [kEigN][kEigN]: u0, u1, u2, l, piv of each lane's factorisation
  double* x;     // [kEigN][kEigN] inverse-iteration iterate of each lane
  double* tau;   // [kEigN] complex
  double* d;     // [kEigN] diagonal of T
  double* e;     // [kEigN] off-diagonal of T
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
// sum over the 32 lanes of a half-wave (every lane gets the sum)
__device__ __forceinline__ cx<double> halfwave_sum(cx<double> v) {
#pragma unroll
  for (int m = 16; m >= 1; m >>= 1) v = v + shfl_xor_cx(v, m);
  return v;
}
// Row sums over the 32 lanes of a half-wave for 8 rows at once (reduce-scatter): on return
// the lane holds the complete sum of row 4*b4 + 2*b3 + b2 (b_k = bit k of its index tc in
// the half-wave) -- 4 + 2 + 1 + 2 shuffles instead of 8 x 5.
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
    a2[r] = keep + shfl_xor_cx(send, 8);
  }
  const cx<double> send = h2 ? a2[0] : a2[1], keep = h2 ? a2[1] : a2[0];
  cx<double> t = keep + shfl_xor_cx(send, 4);
  t = t + shfl_xor_cx(t, 2);
  t = t + shfl_xor_cx(t, 1);
  row = 4 * (int)h4 + 2 * (int)h3 + (int)h2;
  return t;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
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

// number of eigenvalues of T (d, e2 = e^2) below x
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int j = 1; j < n; ++j) {
    q = d[j] - x - e2[j - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

// Phase 1.  A: n x n column-major (lda), upper triangle read (Hermitian completion).
__global__ __launch_bounds__(kEigT) void k_heev_tridiag(const cx<double>* __restrict__ A, int lda, int n,
                                                        EigScratch sc) {
  __shared__ cx<double> vx[2][kEigN];   // x, then v, of the current step (double-buffered)
  __shared__ cx<double> pv[kEigN];      // p = tau A v
  __shared__ cx<double> gpart[kEigG];   // per-row-group partial sums of p^H v
  __shared__ cx<double> tau_s[kEigN];
  const int tid = threadIdx.x;
  const int tr = tid >> 5, tc = tid & 31, lane = tid & 63;

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
      // column block k >> 5 selected with compile-time indices (no dynamic register indexing)
#pragma unroll
      for (int jb = 0; jb < kEigCB; ++jb) {
        if (jb != (k >> 5)) continue;
#pragma unroll
        for (int ia = 0; ia < kEigRA; ++ia) {
          const int i = tr + kEigG * ia;
          if (i > k && i < n) vx[cur][i] = a[ia][jb];
          if (i == k) sc.d[k] = a[ia][jb].x;
        }
      }
    }
    lds_barrier();
    // Householder vector (zlarfg): H^H x = beta e_1, beta real
    if (tid < 64) {
      const cx<double> alpha = vx[cur][k + 1];
      double s = 0.0;
      for (int i = k + 2 + lane; i < n; i += 64) s += vx[cur][i].x * vx[cur][i].x + vx[cur][i].y * vx[cur][i].y;
      const double xn2 = wave_sum(s);
      cx<double> tau = mk<double>(0.0, 0.0), scale = mk<double>(0.0, 0.0);
      double beta = alpha.x;
      if (xn2 != 0.0 || alpha.y != 0.0) {
        const double r = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = alpha.x >= 0.0 ? -r : r;
        tau = mk<double>((beta - alpha.x) / beta, -alpha.y / beta);
        const cx<double> dnm = mk<double>(alpha.x - beta, alpha.y);      // 1 / (alpha - beta)
        const double q = dnm.x * dnm.x + dnm.y * dnm.y;
        scale = mk<double>(dnm.x / q, -dnm.y / q);
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
      for (int jb = 0; jb < kEigCB; ++jb) {
        const int j = tc + 32 * jb;
        vj[jb] = (j > k && j < n) ? vx[cur][j] : mk<double>(0.0, 0.0);
      }
      static_assert(kEigRA == 8, "reduce_scatter8 handles 8 row blocks");
      cx<double> rs[kEigRA];
#pragma unroll
      for (int ia = 0; ia < kEigRA; ++ia) {
        cx<double> sum = mk<double>(0.0, 0.0);
#pragma unroll
        for (int jb = 0; jb < kEigCB; ++jb) sum = sum + cmul(a[ia][jb], vj[jb]);
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
      g = halfwave_sum(g);
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
      for (int jb = 0; jb < kEigCB; ++jb) {
        const int j = tc + 32 * jb;
        const bool in = j > k && j < n;
        vj[jb] = in ? vx[cur][j] : mk<double>(0.0, 0.0);
        wj[jb] = in ? pv[j] + cmul(gam, vj[jb]) : mk<double>(0.0, 0.0);
      }
#pragma unroll
      for (int ia = 0; ia < kEigRA; ++ia) {
        const int i = tr + kEigG * ia;
        const bool in = i > k && i < n;
        const cx<double> vi = in ? vx[cur][i] : mk<double>(0.0, 0.0);
        const cx<double> wi = in ? pv[i] + cmul(gam, vi) : mk<double>(0.0, 0.0);
#pragma unroll
        for (int jb = 0; jb < kEigCB; ++jb) {
          // v_i conj(w_j) + w_i conj(v_j); zero outside the trailing block
          const cx<double> u = cmulc(wj[jb], vi) + cmulc(vj[jb], wi);
          a[ia][jb] = a[ia][jb] - u;
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

// Phases 2 and 3: eigenvalues w (ascending) and eigenvectors Z of T; info[0] = 1 when a
// non-finite value appeared.
__global__ __launch_bounds__(kEigT) void k_heev_trieig(int n, double* __restrict__ w, int* __restrict__ info,
                                                       EigScratch sc) {
  __shared__ double dv[kEigN], ev[kEigN], e2v[kEigN], lam[kEigN];
  __shared__ double scal[5];            // gl, gu, pivmin, tnorm, scale applied to T
  __shared__ int cstart[kEigN];
  __shared__ int maxpos, bad;
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) { maxpos = 0; bad = 0; }
  if (tid < n) {
    dv[tid] = sc.d[tid];
    ev[tid] = tid + 1 < n ? sc.e[tid] : 0.0;
  }
  __syncthreads();

  // ---- phase 2: eigenvalues of T by multisection --------------------------------------
  if (tid < 64) {
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0, dmax = 0.0, emax = 0.0;
    for (int j = lane; j < n; j += 64) {
      const double el = j > 0 ? fabs(ev[j - 1]) : 0.0, er = j + 1 < n ? fabs(ev[j]) : 0.0;
      gl = fmin(gl, dv[j] - el - er);
      gu = fmax(gu, dv[j] + el + er);
      dmax = fmax(dmax, fabs(dv[j]));
      if (j + 1 < n) { emax2 = fmax(emax2, ev[j] * ev[j]); emax = fmax(emax, er); e2v[j] = ev[j] * ev[j]; }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      gl = fmin(gl, __shfl_xor(gl, m));
      gu = fmax(gu, __shfl_xor(gu, m));
      emax2 = fmax(emax2, __shfl_xor(emax2, m));
      dmax = fmax(dmax, __shfl_xor(dmax, m));
      emax = fmax(emax, __shfl_xor(emax, m));
    }
    // T = 0: eigenvalues 0, eigenvectors e_i.  Otherwise T is scaled to unit norm when
    // its norm is far from 1 (overflow-free inverse iteration), and lambda scaled back.
    const double tn0 = dmax + 2.0 * emax;
    const double sc_ = (tn0 > 0.0 && (tn0 < 1e-100 || tn0 > 1e100)) ? 1.0 / tn0 : 1.0;
    if (sc_ != 1.0) {
      for (int j = lane; j < n; j += 64) {
        dv[j] *= sc_;
        if (j + 1 < n) { ev[j] *= sc_; e2v[j] = ev[j] * ev[j]; }
      }
      gl *= sc_; gu *= sc_; emax2 *= sc_ * sc_; dmax *= sc_; emax *= sc_;
    }
    if (lane == 0) {
      scal[4] = sc_;
      const double tn = fmax(fabs(gl), fabs(gu));
      const double pad = 2.0 * DBL_EPSILON * tn * n + DBL_MIN;
      scal[0] = gl - pad;
      scal[1] = gu + pad;
      scal[2] = fmax(emax2, 1.0) * (DBL_MIN / DBL_EPSILON);
      scal[3] = dmax + 2.0 * emax;
    }
  }
  __syncthreads();
  {
    const int i = tid / kEigSamples, s = tid % kEigSamples;
    if (i < n) {
      double lo = scal[0], hi = scal[1];
      const double pivmin = scal[2];
      for (int round = 0; round < 64; ++round) {
        if (!(hi - lo > 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + DBL_MIN)) break;
        const double x = lo + (hi - lo) * (double)(s + 1) / (double)(kEigSamples + 1);
        const int c = sturm_count(dv, e2v, n, x, pivmin);
        double nlo = c <= i ? x : lo, nhi = c > i ? x : hi;
#pragma unroll
        for (int m = 1; m < kEigSamples; m <<= 1) {
          nlo = fmax(nlo, __shfl_xor(nlo, m));
          nhi = fmin(nhi, __shfl_xor(nhi, m));
        }
        if (nlo == lo && nhi == hi) break;
        lo = nlo;
        hi = nhi;
      }
      if (s == 0) lam[i] = 0.5 * (lo + hi);
    }
  }
  __syncthreads();

  // ---- phase 3: eigenvectors of T by inverse iteration --------------------------------
  const double tnorm = scal[3];
  // clusters for Gram-Schmidt: gaps below 1e-6 ||T|| (exact and near multiplets); the
  // overlaps between vectors of wider gaps, at most ~eps ||T|| / gap, are removed by the
  // Loewdin step 3c
  const double ortol = 1e-6 * tnorm, pivtol = DBL_EPSILON * tnorm;
  if (tnorm == 0.0) {
    // T = 0 (A = 0): identity eigenvectors
    if (tid < n)
      for (int k = 0; k < n; ++k) sc.Z[(size_t)tid * kEigN + k] = k == tid ? 1.0 : 0.0;
    if (tid < n) w[tid] = 0.0;
    if (tid == 0) info[0] = 0;
    return;
  }
  // 3a: inverse iteration, one lane per eigenvalue, all eigenvalues at once
  double* u0 = sc.lu;
  double* u1 = sc.lu + kEigN * kEigN;
  double* u2 = sc.lu + 2 * kEigN * kEigN;
  double* ll = sc.lu + 3 * kEigN * kEigN;
  double* pv_ = sc.lu + 4 * kEigN * kEigN;
  double* x = sc.x;
  if (tid < n) {
    const int i = tid;
    const double lmb = lam[i];
    // T - lambda I = P L U, row interchanges where the subdiagonal dominates; the factors
    // of lane i live at [k][i] (consecutive lanes, consecutive addresses)
    double cd = dv[0] - lmb, cs = n > 1 ? ev[0] : 0.0;
    for (int k = 0; k + 1 < n; ++k) {
      const double sub = ev[k], nd = dv[k + 1] - lmb, ns = k + 2 < n ? ev[k + 1] : 0.0;
      const size_t o = (size_t)k * kEigN + i;
      if (fabs(cd) >= fabs(sub)) {
        const double m = cd != 0.0 ? sub / cd : 0.0;
        u0[o] = cd; u1[o] = cs; u2[o] = 0.0; ll[o] = m; pv_[o] = 0.0;
        cd = nd - m * cs;
        cs = ns;
      } else {
        const double m = cd / sub;
        u0[o] = sub; u1[o] = nd; u2[o] = ns; ll[o] = m; pv_[o] = 1.0;
        const double t = cs - m * nd;
        cs = -m * ns;
        cd = t;
      }
    }
    u0[(size_t)(n - 1) * kEigN + i] = cd;
    // start vector b_k = cos(k (1 + 0.37 i) + i / 2), then three solves (T - lambda I) x = b / |b|
    double nrm = 0.0;
    for (int k = 0; k < n; ++k) {
      const double b = cos((double)k * (1.0 + 0.37 * i) + 0.5 * i);
      x[(size_t)k * kEigN + i] = b;
      nrm += b * b;
    }
    double inv = 1.0 / sqrt(nrm);
    for (int it = 0; it < 3; ++it) {
      double yk = x[i] * inv;
#pragma unroll 4
      for (int k = 0; k + 1 < n; ++k) {
        const size_t o = (size_t)k * kEigN + i;
        double yn = x[o + kEigN] * inv;
        if (pv_[o] != 0.0) { const double t = yk; yk = yn; yn = t; }
        x[o] = yk;
        yk = yn - ll[o] * yk;
      }
      x[(size_t)(n - 1) * kEigN + i] = yk;
      double x1 = 0.0, x2 = 0.0;
      nrm = 0.0;
#pragma unroll 4
      for (int k = n - 1; k >= 0; --k) {
        const size_t o = (size_t)k * kEigN + i;
        double u = u0[o];
        if (fabs(u) < pivtol) u = u < 0.0 ? -pivtol : pivtol;
        double rhs = x[o];
        if (k + 1 < n) rhs -= u1[o] * x1;
        if (k + 2 < n) rhs -= u2[o] * x2;
        const double xv = rhs / u;
        x[o] = xv;
        x2 = x1;
        x1 = xv;
        nrm += xv * xv;
      }
      inv = 1.0 / sqrt(nrm);
    }
    // eigenvector i stored contiguously: Z[i * kEigN + k]
    for (int k = 0; k < n; ++k) sc.Z[(size_t)i * kEigN + k] = x[(size_t)k * kEigN + i] * inv;
    if (!(nrm > 0.0 && nrm <= DBL_MAX) || !(lmb == lmb)) atomicOr(&bad, 1);
    // cluster: the run of eigenvalues with consecutive gaps <= ortol that contains i
    int st = i;
    while (st > 0 && lam[st] - lam[st - 1] <= ortol) --st;
    cstart[i] = st;
    atomicMax(&maxpos, i - st);
  }
  __syncthreads();
  if (tid == 0) e2v[0] = 0.0;   // Sturm counts are done; e2v[0] holds max |E| in step 3c
  // 3b: orthonormalise every cluster member against the earlier ones (modified
  // Gram-Schmidt, two passes), one half-wave per member, members of the same rank in
  // parallel; a lane holds components lane + 32c
  {
    const int hw = tid >> 5, hl = tid & 31;
    const int rounds = maxpos;
    for (int r = 1; r <= rounds; ++r) {
      for (int i = hw; i < n; i += kEigG) {
        const int st = cstart[i];
        if (i - st != r) continue;
        double xv[kEigCB];
#pragma unroll
        for (int c = 0; c < kEigCB; ++c) {
          const int k = hl + 32 * c;
          xv[c] = k < n ? sc.Z[(size_t)i * kEigN + k] : 0.0;
        }
        for (int pass = 0; pass < 2; ++pass) {
          for (int j = st; j < i; ++j) {
            double zj[kEigCB], dot = 0.0;
#pragma unroll
            for (int c = 0; c < kEigCB; ++c) {
              const int k = hl + 32 * c;
              zj[c] = k < n ? sc.Z[(size_t)j * kEigN + k] : 0.0;
              dot += zj[c] * xv[c];
            }
#pragma unroll
            for (int m = 16; m >= 1; m >>= 1) dot += __shfl_xor(dot, m);
#pragma unroll
            for (int c = 0; c < kEigCB; ++c) xv[c] -= dot * zj[c];
          }
        }
        double nn = 0.0;
#pragma unroll
        for (int c = 0; c < kEigCB; ++c) nn += xv[c] * xv[c];
#pragma unroll
        for (int m = 16; m >= 1; m >>= 1) nn += __shfl_xor(nn, m);
        if (!(nn > 0.0 && nn <= DBL_MAX) && hl == 0) atomicOr(&bad, 1);
        const double s = 1.0 / sqrt(nn);
#pragma unroll
        for (int c = 0; c < kEigCB; ++c) {
          const int k = hl + 32 * c;
          if (k < n) sc.Z[(size_t)i * kEigN + k] = xv[c] * s;
        }
      }
      __syncthreads();
    }
  }

  // 3c: Loewdin orthonormalisation Z <- Z - 1/2 (Z Z^T - I) Z, repeated while the largest
  // overlap exceeds 1e-14 (at most 3 times; first order, so each pass squares the error)
  {
    extern __shared__ __align__(16) unsigned char tsm[];
    double* zs = reinterpret_cast<double*>(tsm);           // [kEigN][kEigN + 1]
    constexpr int ZS = kEigN + 1;
    double* E = sc.lu;                                     // [kEigN][kEigN] in global scratch
    const int ti = tid >> 5, tj = tid & 31;
    for (int pass = 0; pass < 3; ++pass) {
      for (int t = tid; t < n * n; t += kEigT) {
        const int i = t / n, k = t - i * n;
        zs[i * ZS + k] = sc.Z[(size_t)i * kEigN + k];
      }
      if (tid == 0) maxpos = 0;
      __syncthreads();
      // E[i][j] for rows i = ti + 16a, columns j = tj + 32b
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
      double emax = 0.0;
#pragma unroll
      for (int a = 0; a < kEigRA; ++a)
#pragma unroll
        for (int b = 0; b < kEigCB; ++b) {
          const int i = ti + kEigG * a, j = tj + 32 * b;
          if (i < n && j < n) {
            const double e = acc[a][b] - (i == j ? 1.0 : 0.0);
            E[(size_t)i * kEigN + j] = e;
            emax = fmax(emax, fabs(e));
          }
        }
      // max |E| as an integer bit pattern (non-negative doubles order like their bits)
      const unsigned long long eb = __double_as_longlong(emax);
      atomicMax(reinterpret_cast<unsigned long long*>(&e2v[0]), eb);
      __syncthreads();
      const double emx = __longlong_as_double((long long)*reinterpret_cast<unsigned long long*>(&e2v[0]));
      if (!(emx > 1e-14)) break;
      if (!(emx < 0.1)) { if (tid == 0) atomicOr(&bad, 1); break; }
      // Z'[i][k] = Z[i][k] - 1/2 sum_j E[i][j] Z[j][k], rows i = ti + 16a, components k = tj + 32b
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
#pragma unroll
      for (int a = 0; a < kEigRA; ++a)
#pragma unroll
        for (int b = 0; b < kEigCB; ++b) {
          const int i = ti + kEigG * a, k = tj + 32 * b;
          if (i < n && k < n) sc.Z[(size_t)i * kEigN + k] = zs[i * ZS + k] - 0.5 * acc[a][b];
        }
      __syncthreads();
      if (tid == 0) e2v[0] = 0.0;
      __syncthreads();
    }
  }

  if (tid < n) w[tid] = lam[tid] / scal[4];
  if (tid == 0) info[0] = bad;
}

// Phase 4: A <- Q Z (the eigenvectors of A, column-major, lda).  The Householder vectors
// are staged in LDS first (packed lower triangle, n (n-1) / 2 complex), so the chain of
// n - 1 reflections reads no global memory.
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
  const int rr = tid & 31, cc = tid >> 5;
  cx<double> y[kEigCB][kEigRA];
#pragma unroll
  for (int ia = 0; ia < kEigCB; ++ia)
#pragma unroll
    for (int jb = 0; jb < kEigRA; ++jb) {
      const int i = rr + 32 * ia, j = cc + kEigG * jb;
      y[ia][jb] = mk<double>((i < n && j < n) ? sc.Z[(size_t)j * kEigN + i] : 0.0, 0.0);
    }
  for (int k = n - 2; k >= 0; --k) {
    const cx<double> tk = tp[k];
    const int o = vpack_off(k, n) - (k + 1);
    cx<double> v[kEigCB];
#pragma unroll
    for (int ia = 0; ia < kEigCB; ++ia) {
      const int i = rr + 32 * ia;
      v[ia] = (i > k && i < n) ? vp[o + i] : mk<double>(0.0, 0.0);
    }
#pragma unroll
    for (int jb = 0; jb < kEigRA; ++jb) {
      cx<double> s = mk<double>(0.0, 0.0);
#pragma unroll
      for (int ia = 0; ia < kEigCB; ++ia) s = s + cmulc(v[ia], y[ia][jb]);
      s = cmul(tk, halfwave_sum(s));
#pragma unroll
      for (int ia = 0; ia < kEigCB; ++ia) y[ia][jb] = y[ia][jb] - cmul(v[ia], s);
    }
  }
#pragma unroll
  for (int ia = 0; ia < kEigCB; ++ia)
#pragma unroll
    for (int jb = 0; jb < kEigRA; ++jb) {
      const int i = rr + 32 * ia, j = cc + kEigG * jb;
      if (i < n && j < n) A[i + (size_t)j * lda] = y[ia][jb];
    }
}

size_t heev_small_scratch_doubles() { return (size_t)kEigN * kEigN * (2 + 1 + 5 + 1) + 4 * kEigN; }

hipError_t launch_heev_small(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                             hipStream_t s) {
  if (n < 1 || n > kEigN || lda < n) return hipErrorInvalidValue;
  EigScratch sc;
  sc.V = scratch;
  sc.Z = scratch + 2 * kEigN * kEigN;
  sc.lu = sc.Z + kEigN * kEigN;
  sc.x = sc.lu + 5 * kEigN * kEigN;
  sc.tau = sc.x + kEigN * kEigN;
  sc.d = sc.tau + 2 * kEigN;
  sc.e = sc.d + kEigN;
  hipLaunchKernelGGL(k_heev_tridiag, dim3(1), dim3(kEigT), 0, s, A, lda, n, sc);
  constexpr size_t tsm = sizeof(double) * (size_t)kEigN * (kEigN + 1);
  static bool tattr = false;
  if (!tattr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_trieig),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)tsm);
    if (e != hipSuccess) return e;
    tattr = true;
  }
  hipLaunchKernelGGL(k_heev_trieig, dim3(1), dim3(kEigT), tsm, s, n, w, info, sc);
  constexpr size_t bsm = sizeof(cx<double>) * ((size_t)kEigN * (kEigN - 1) / 2 + kEigN);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_backtr),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bsm);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_heev_backtr, dim3(1), dim3(kEigT), bsm, s, A, lda, n, sc);
  return hipGetLastError();
}

}  // namespace pwh
