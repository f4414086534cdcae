// pwh_gemm.hip -- complex GEMMs on the gfx950 matrix cores.
//
// Two shapes cover the nonlocal projectors and the subspace matrices:
//
//   proj   C[m][n] = sum_k conj(A[m][k]) * B[n][k]      (reduction over plane waves)
//          - nonlocal:  A = beta[(a,l)][k] = EIGR[a][k]*PKG_A[l][k] (stored once per
//            system), B = psi[b][k]  ->  <beta_(a,l)|psi_b>
//            (ApplyHam_block.cpp:88-117: conj(EIGR)*PKG_A then zdotu with vin)
//          - subspace:  S = X^H X, T = X^H (HX)  (ChebySCFLoop.cpp:296)
//          split over K into deterministic partial sums, then a reduction that can
//          apply the per-projector D = WNL(l) scale.
//   expand out[n][k] (+)= sum_j A[j][k] * Q[j][n]      (reduction over a small index)
//          - nonlocal:  A = beta (stored), Q = D <beta|psi>  ->  sum beta D <beta|psi>
//            (ApplyHam_block.cpp:118-163)
//          - rotation:  X <- X Q (ChebySCFLoop.cpp:304-306, CholeskyQR's X R^-1)
//
// Precisions:
//   FP64   v_mfma_f64_16x16x4_f64, three real products per complex one (3M)
//   FP32   v_mfma_f32_16x16x4_f32, three real products per complex one (3M)
//   BF16X3 / BF16X6  v_mfma_f32_16x16x32_bf16 on two- / three-term bf16 splits of the FP32
//          operands (3 / 6 split products per real product), FP32 accumulation
// Block tile 64x64 complex, 4 waves of 32x32, LDS staging.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "pwh_blas.h"
#include "pwh_internal.h"

namespace pwh {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T> struct AccT;
template <> struct AccT<double> { using type = d4v; };
template <> struct AccT<float> { using type = f4v; };

__device__ __forceinline__ d4v mfma4(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 result held in accumulator register r of a lane in row-group lq
// (lq = lane / 16; the column is lane % 16).  The f32 and bf16 instructions give each
// lane four consecutive rows; the f64 instruction interleaves them (register r holds
// row 4r + lq).
template <typename T> __device__ __forceinline__ int acc_row(int lq, int r);
template <> __device__ __forceinline__ int acc_row<float>(int lq, int r) { return 4 * lq + r; }
template <> __device__ __forceinline__ int acc_row<double>(int lq, int r) { return 4 * r + lq; }

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

constexpr int BT = 64;    // block tile (both output dims)
// workgroups per CU the register allocation must allow: the FP64 kernels hold three FP64
// accumulator sets (3M) and would otherwise fall to one wave per SIMD
template <typename T> constexpr int gemm_min_blocks() { return sizeof(T) == 8 ? 2 : 1; }
constexpr int LPAD = 4;
constexpr int GT = 256;   // threads per block

// Row loader of a k-contiguous matrix A[m][k] (lda).  The BETA flag is kept in the kernel
// signatures for the tile shapes; only the plain form is instantiated.
template <typename T, bool BETA>
struct RowLoader {
  const cx<T>* A; int lda;
  __device__ __forceinline__ cx<T> operator()(int m, int k) const { return A[(size_t)m * lda + k]; }
};

// beta[(a, l)][k] generated from the factorised phases (PhaseFactor.cpp:41-77 forms EIGR
// as the same product) and PKG_A: three table reads, two complex products and a scale in
// place of one read of the stored row
template <>
struct RowLoader<float, true> {
  const cx<float>* A; int lda;          // unused (the stored rows)
  const cx<float>* ei;
  const int* mi;
  const float* pkg;
  int ldt, lm, ldp, slot0;
  __device__ __forceinline__ cx<float> operator()(int m, int k) const {
    const int a = m / lm, l = m - a * lm;
    const int p = mi[k];
    const cx<float>* t = ei + (size_t)(slot0 + a) * 3 * ldt;
    const cx<float> e = (t[p & 1023] * t[ldt + ((p >> 10) & 1023)]) * t[2 * ldt + (p >> 20)];
    const float s = pkg[(size_t)l * ldp + k];
    return mk<float>(e.x * s, e.y * s);
  }
};

// ---------------------------------------------------------------------------
// proj: part[s][m][n] = sum_{k in chunk s} conj(A[m][k]) B[n][k]
// ---------------------------------------------------------------------------
// LDS of one block: four planes (Re/Im of both operands) of KS rows of BT + LPAD
template <typename T, int KS>
constexpr size_t gemm_smem() { return 4 * sizeof(T) * (size_t)KS * (BT + LPAD); }

// Block L of the projection grid (the kernel body, also run by the grouped kernel).
// UPPER: only the tiles that reach the upper triangle (m <= n) are computed; the others
// exit (a Hermitian product whose lower triangle is not read).
template <typename T, bool BETA, int KS, bool UPPER = false>
__device__ __forceinline__ void gemm_proj_block(RowLoader<T, BETA> ld, const cx<T>* __restrict__ B, int ldb,
                                                int M, int N, int K, int kchunk, cx<T>* __restrict__ part,
                                                const int L) {
  using acc_t = typename AccT<T>::type;
  static_assert(KS % 16 == 0, "reduction depth per stage: multiple of 16");
  extern __shared__ __align__(16) unsigned char gsm[];
  typedef T Row[BT + LPAD];
  Row* sAr = reinterpret_cast<Row*>(gsm);
  Row* sAi = sAr + KS;
  Row* sBr = sAi + KS;
  Row* sBi = sBr + KS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // UPPER (M == N): the grid enumerates only the tiles on or above the diagonal
  const int tiles_n = (N + BT - 1) / BT;
  const int tiles = UPPER ? tiles_n * (tiles_n + 1) / 2 : ((M + BT - 1) / BT) * tiles_n;
  // XCD-aware order: workgroup L of the 1D grid runs on XCD L % 8 (round-robin dispatch).
  // The tiles of one K-split read the same chunk of both operands, so they get the same
  // L % 8 and neighbouring L: the re-reads hit that XCD's L2.  Splits past the last exit.
  // With a split count that is not a multiple of 8 the plain order (tile fastest) is used.
  const int q = L >> 3;
  const bool xcd = ((K + kchunk - 1) / kchunk) % 8 == 0;
  const int tile = xcd ? q % tiles : L % tiles;
  const int split = xcd ? (L & 7) + 8 * (q / tiles) : L / tiles;
  if (split * kchunk >= K) return;
  int mt = tile / tiles_n, nt = tile % tiles_n;
  if (UPPER) {
    // upper tile t, rows of tiles in order: row r holds tiles_n - r tiles (r .. tiles_n - 1)
    int r = 0, t = tile;
    while (t >= tiles_n - r) {
      t -= tiles_n - r;
      ++r;
    }
    mt = r;
    nt = r + t;
  }
  const int m0 = mt * BT, n0 = nt * BT;
  const int ks = split * kchunk, ke = min(K, ks + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  // three real products per complex one (3M): cr, c2, ci accumulate S1 = sum ar br,
  // S2 = sum ai bi and S3 = sum of the (sum/difference) product
  acc_t cr[2][2], ci[2][2], c2[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0}; ci[a][b] = acc_t{0, 0, 0, 0}; c2[a][b] = acc_t{0, 0, 0, 0};
    }

  // register prefetch of the next K tile overlaps its global loads with the MFMAs
  // element e of a stage: 16 consecutive k of one row per 16 lanes (row r = (e/16) % 64,
  // k = 16 (e/1024) + e%16): 128-byte row segments, conflict-free transposed LDS stores
  constexpr int PER = BT * KS / GT;
  cx<T> pa[PER], pb[PER];
  auto slot = [](int e, int& kk, int& rr) { kk = (e & 15) + 16 * (e >> 10); rr = (e >> 4) & (BT - 1); };
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * GT;
      int kk, rr;
      slot(e, kk, rr);
      const int k = k0 + kk;
      pa[q] = mk<T>(0, 0);
      pb[q] = mk<T>(0, 0);
      if (k < ke) {
        if (m0 + rr < M) pa[q] = ld(m0 + rr, k);
        if (n0 + rr < N) pb[q] = B[(size_t)(n0 + rr) * ldb + k];
      }
    }
  };
  fetch(ks);
  for (int k0 = ks; k0 < ke; k0 += KS) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * GT;
      int kk, rr;
      slot(e, kk, rr);
      sAr[kk][rr] = pa[q].x; sAi[kk][rr] = pa[q].y;
      sBr[kk][rr] = pb[q].x; sBi[kk][rr] = pb[q].y;
    }
    __syncthreads();
    if (k0 + KS < ke) fetch(k0 + KS);
#pragma unroll
    for (int kc = 0; kc < KS; kc += 16) {
      // substep sub reduces over k = kc + 4*lq + sub: lane quarters then read rows 4 apart,
      // 4*(BT+LPAD) = 16 banks, so the four quarters of a wave never share a bank
#pragma unroll
      for (int sub = 0; sub < 4; ++sub) {
        T ar[2], ai[2], br[2], bi[2];
        const int kr = kc + 4 * lq + sub;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ar[t] = sAr[kr][wm + t * 16 + li]; ai[t] = sAi[kr][wm + t * 16 + li];
          br[t] = sBr[kr][wn + t * 16 + li]; bi[t] = sBi[kr][wn + t * 16 + li];
        }
        // conj(a) b: S1 = ar br, S2 = ai bi, S3 = (ar - ai)(br + bi); Re = S1 + S2,
        // Im = S3 - S1 + S2.  Consecutive MFMAs never feed each other.
        T as[2], bs[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) { as[t] = ar[t] - ai[t]; bs[t] = br[t] + bi[t]; }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            cr[a][b] = mfma4(ar[a], br[b], cr[a][b]);
            c2[a][b] = mfma4(ai[a], bi[b], c2[a][b]);
            ci[a][b] = mfma4(as[a], bs[b], ci[a][b]);
          }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const T s1 = cr[a][b][r], s2 = c2[a][b][r], s3 = ci[a][b][r];
        cr[a][b][r] = s1 + s2;
        ci[a][b][r] = (s3 - s1) + s2;
      }
  cx<T>* P = part + (size_t)split * M * N;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + acc_row<T>(lq, r), n = n0 + wn + b * 16 + li;
        if (m < M && n < N) P[(size_t)m * N + n] = mk<T>(cr[a][b][r], ci[a][b][r]);
      }
}

template <typename T, bool BETA, int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<T>()) void k_gemm_proj(RowLoader<T, BETA> ld, const cx<T>* __restrict__ B,
                                                  int ldb, int M, int N, int K, int kchunk,
                                                  cx<T>* __restrict__ part) {
  gemm_proj_block<T, BETA, KS>(ld, B, ldb, M, N, K, kchunk, part, blockIdx.x);
}

template <int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<double>()) void k_gemm_gram_upper(RowLoader<double, false> ld, const cx<double>* __restrict__ B,
                                                        int ldb, int M, int N, int K, int kchunk,
                                                        cx<double>* __restrict__ part) {
  gemm_proj_block<double, false, KS, true>(ld, B, ldb, M, N, K, kchunk, part, blockIdx.x);
}

// C = scale(m) * sum_s part[s][m][n], stored C[m*ldc + n] (row-major) or C[m + n*ldc]
// (column-major, the LAPACK layout handed to rocSOLVER); scale(m) = wl[m % lm] when wl != null.
// One output per thread: eight independent partial sums (split s goes to sum s % 8) keep
// eight loads in flight, then a fixed tree adds them (deterministic).
template <typename T, typename TO>
__device__ __forceinline__ void reduce_parts_block(const cx<T>* __restrict__ part, int S, int M, int N,
                                                   cx<TO>* __restrict__ C, int ldc, const T* __restrict__ wl,
                                                   int lm, int colmajor, const int blk) {
  const int idx = blk * 256 + threadIdx.x;
  const size_t MN = (size_t)M * N;
  if (idx >= M * N) return;
  TO ar[8], ai[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) { ar[j] = 0; ai[j] = 0; }
  int s = 0;
  for (; s + 8 <= S; s += 8) {
    cx<T> v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = part[(size_t)(s + j) * MN + idx];
#pragma unroll
    for (int j = 0; j < 8; ++j) { ar[j] += (TO)v[j].x; ai[j] += (TO)v[j].y; }
  }
  for (int j = 0; s < S; ++s, ++j) {
    const cx<T> v = part[(size_t)s * MN + idx];
    ar[j] += (TO)v.x; ai[j] += (TO)v.y;
  }
  TO tr = ((ar[0] + ar[1]) + (ar[2] + ar[3])) + ((ar[4] + ar[5]) + (ar[6] + ar[7]));
  TO ti = ((ai[0] + ai[1]) + (ai[2] + ai[3])) + ((ai[4] + ai[5]) + (ai[6] + ai[7]));
  const int m = idx / N, n = idx - m * N;
  if (wl) { const TO w = (TO)wl[m % lm]; tr *= w; ti *= w; }
  if (colmajor) C[m + (size_t)n * ldc] = mk<TO>(tr, ti);
  else C[(size_t)m * ldc + n] = mk<TO>(tr, ti);
}

template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_reduce_parts(const cx<T>* __restrict__ part, int S, int M,
                                                      int N, cx<TO>* __restrict__ C, int ldc,
                                                      const T* __restrict__ wl, int lm, int colmajor) {
  reduce_parts_block<T, TO>(part, S, M, N, C, ldc, wl, lm, colmajor, blockIdx.x);
}

// ---------------------------------------------------------------------------
// expand: out[n][k] (+)= sum_j A[j][k] Q[j][n]
// ---------------------------------------------------------------------------
// TRI: Q is upper triangular (Q[j][n] = 0 for j > n), so the tile of columns n0.. reduces
// over j < n0 + BT only (the skipped products are exact zeros: the result is unchanged)
template <typename T, bool BETA, bool QCM, int KS, bool TRI = false>
__device__ __forceinline__ void gemm_expand_block(RowLoader<T, BETA> ld, const cx<T>* __restrict__ Q, int ldq,
                                                  int J, int N, int K, cx<T>* __restrict__ out, int ldo,
                                                  int accumulate, const int L) {
  using acc_t = typename AccT<T>::type;
  static_assert(KS % 16 == 0, "reduction depth per stage: multiple of 16");
  extern __shared__ __align__(16) unsigned char gsm[];
  typedef T Row[BT + LPAD];
  Row* sAr = reinterpret_cast<Row*>(gsm);
  Row* sAi = sAr + KS;
  Row* sQr = sAi + KS;
  Row* sQi = sQr + KS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware order (as in k_gemm_proj): the band tiles of one plane-wave tile read the
  // same rows of A, so they share L % 8 and run close together; tiles past K exit.
  const int tiles_n = (N + BT - 1) / BT;
  const int q = L >> 3;
  const int k0 = ((L & 7) + 8 * (q / tiles_n)) * BT, n0 = (q % tiles_n) * BT;
  if (k0 >= K) return;
  if (TRI) J = min(J, n0 + BT);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  // three real products per complex one (3M): cr, c2, ci accumulate S1 = sum ar br,
  // S2 = sum ai bi and S3 = sum of the (sum/difference) product
  acc_t cr[2][2], ci[2][2], c2[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0}; ci[a][b] = acc_t{0, 0, 0, 0}; c2[a][b] = acc_t{0, 0, 0, 0};
    }

  constexpr int PER = BT * KS / GT;
  cx<T> pa[PER], pq[PER];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * GT;
      const int kk = e % BT, jj = e / BT;
      const int j = j0 + jj;
      pa[q] = mk<T>(0, 0);
      pq[q] = mk<T>(0, 0);
      if (j < J) {
        if (k0 + kk < K) pa[q] = ld(j, k0 + kk);
        if (n0 + kk < N) pq[q] = QCM ? Q[j + (size_t)(n0 + kk) * ldq] : Q[(size_t)j * ldq + n0 + kk];
      }
    }
  };
  fetch(0);
  for (int j0 = 0; j0 < J; j0 += KS) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * GT;
      const int kk = e % BT, jj = e / BT;
      sAr[jj][kk] = pa[q].x; sAi[jj][kk] = pa[q].y;
      sQr[jj][kk] = pq[q].x; sQi[jj][kk] = pq[q].y;
    }
    __syncthreads();
    if (j0 + KS < J) fetch(j0 + KS);
#pragma unroll
    for (int kc = 0; kc < KS; kc += 16) {
      // reduction index j = kc + 4*lq + sub (conflict-free rows, as in k_gemm_proj)
#pragma unroll
      for (int sub = 0; sub < 4; ++sub) {
        T ar[2], ai[2], br[2], bi[2];
        const int jr = kc + 4 * lq + sub;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ar[t] = sAr[jr][wm + t * 16 + li]; ai[t] = sAi[jr][wm + t * 16 + li];
          br[t] = sQr[jr][wn + t * 16 + li]; bi[t] = sQi[jr][wn + t * 16 + li];
        }
        // a q: S1 = ar qr, S2 = ai qi, S3 = (ar + ai)(qr + qi); Re = S1 - S2,
        // Im = S3 - S1 - S2
        T as[2], bs[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) { as[t] = ar[t] + ai[t]; bs[t] = br[t] + bi[t]; }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            cr[a][b] = mfma4(ar[a], br[b], cr[a][b]);
            c2[a][b] = mfma4(ai[a], bi[b], c2[a][b]);
            ci[a][b] = mfma4(as[a], bs[b], ci[a][b]);
          }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const T s1 = cr[a][b][r], s2 = c2[a][b][r], s3 = ci[a][b][r];
        cr[a][b][r] = s1 - s2;
        ci[a][b][r] = (s3 - s1) - s2;
      }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + wm + a * 16 + acc_row<T>(lq, r), n = n0 + wn + b * 16 + li;
        if (k < K && n < N) {
          cx<T> v = mk<T>(cr[a][b][r], ci[a][b][r]);
          cx<T>* o = out + (size_t)n * ldo + k;
          if (accumulate) v = v + *o;
          *o = v;
        }
      }
}

template <typename T, bool BETA, bool QCM, int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<T>()) void k_gemm_expand(RowLoader<T, BETA> ld, const cx<T>* __restrict__ Q,
                                                    int ldq, int J, int N, int K,
                                                    cx<T>* __restrict__ out, int ldo, int accumulate) {
  gemm_expand_block<T, BETA, QCM, KS>(ld, Q, ldq, J, N, K, out, ldo, accumulate, blockIdx.x);
}

template <int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<double>()) void k_gemm_nn_tri(RowLoader<double, false> ld, const cx<double>* __restrict__ Q,
                                                    int ldq, int J, int N, int K, cx<double>* __restrict__ out,
                                                    int ldo) {
  gemm_expand_block<double, false, true, KS, true>(ld, Q, ldq, J, N, K, out, ldo, 0, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Grouped projector GEMMs: block L of a launch belongs to the problem whose first block is
// the largest one <= L, and runs the one-problem body with its local index.
// ---------------------------------------------------------------------------
template <typename P, typename F>
__device__ __forceinline__ int group_problem(const P* __restrict__ g, int n, int L, F first) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first(g[mid]) <= L) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<T>()) void k_gemm_proj_group(const GroupProj<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupProj<T>& p = g[group_problem(g, n, L, [](const GroupProj<T>& x) { return x.blk0; })];
  RowLoader<T, false> ld{p.A, p.lda};
  gemm_proj_block<T, false, KS>(ld, p.B, p.ldb, p.M, p.N, p.K, p.kchunk, p.part, L - p.blk0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_reduce_group(const GroupProj<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupProj<T>& p = g[group_problem(g, n, L, [](const GroupProj<T>& x) { return x.rblk0; })];
  reduce_parts_block<T, T>(p.part, p.S, p.M, p.N, p.C, p.ldc, p.wl, p.lm, 0, L - p.rblk0);
}

template <typename T, int KS>
__global__ __launch_bounds__(GT, gemm_min_blocks<T>()) void k_gemm_expand_group(const GroupExpand<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupExpand<T>& p = g[group_problem(g, n, L, [](const GroupExpand<T>& x) { return x.blk0; })];
  RowLoader<T, false> ld{p.A, p.lda};
  gemm_expand_block<T, false, false, KS>(ld, p.Q, p.ldq, p.J, p.N, p.K, p.out, p.ldo, p.accumulate, L - p.blk0);
}

// ---------------------------------------------------------------------------
// BF16 split GEMMs on the gfx950 16x16x32 BF16 MFMA
// ---------------------------------------------------------------------------
// Each FP32 operand x is split into NS bf16 terms, x ~ x_0 + x_1 (+ x_2), each term the
// rounding of what the earlier ones leave, and a real product keeps the split products of
// order <= 2^-8 (NS = 2: x0 y0 + x0 y1 + x1 y0, ~16 significant bits) or <= 2^-16 (NS = 3:
// the six products with i + j <= 2, ~24 bits, FP32 level), accumulated in FP32 (the
// construction of Henry, Tang and Heinecke, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations", ARITH 2019).  Operands are
// split once, when a block stages its tile into LDS, as planes [component][term][row][k]
// with the reduction index contiguous, so a lane reads its 8-element MFMA fragment as one
// 16-byte word.  Complex products take four real ones (Re and Im of each operand; the
// negated imaginary fragment is a sign flip of the bf16 words).  Block tile 64 x 64, four
// waves of 32 x 32, 32 reduction steps per stage.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f4v mfma32bf(bf8v a, bf8v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr int BKS = 32;          // reduction depth of one stage (one MFMA)
constexpr int BKP = BKS + 8;     // LDS row length in bf16 (80 bytes, 16-byte aligned)

template <int NS>
constexpr size_t bf_smem() { return sizeof(unsigned short) * 2 * 2 * NS * (size_t)BT * BKP; }

// plane (operand op, component c, term i): BT rows of BKP bf16
template <int NS>
__device__ __forceinline__ unsigned short* bf_plane(unsigned short* base, int op, int c, int i) {
  return base + (size_t)((op * 2 + c) * NS + i) * BT * BKP;
}

// Splits 8 consecutive values (of one component) and stores the NS packed 16-byte words at
// row `row`, reduction offset k8 of the operand's planes.
// A pair is rounded to nearest-even by one v_cvt_pk_bf16_f32; the remainder r - rn(r) is
// exact in FP32.
template <int NS>
__device__ __forceinline__ void bf_store8(unsigned short* base, int op, int c, int row, int k8, const float (&x)[8]) {
  unsigned int w[NS][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float r0 = x[2 * q], r1 = x[2 * q + 1];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const unsigned int u = __builtin_bit_cast(unsigned int, __builtin_convertvector(f2v{r0, r1}, bf2v));
      w[i][q] = u;
      if (i + 1 < NS) {
        r0 -= __uint_as_float(u << 16);
        r1 -= __uint_as_float(u & 0xffff0000u);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i)
    *reinterpret_cast<u4v*>(bf_plane<NS>(base, op, c, i) + row * BKP + k8) = u4v{w[i][0], w[i][1], w[i][2], w[i][3]};
}

// The same for 4 consecutive values: NS 8-byte words at row `row`, reduction offset k4.
template <int NS>
__device__ __forceinline__ void bf_store4(unsigned short* base, int op, int c, int row, int k4, const float (&x)[4]) {
  unsigned int w[NS][2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    float r0 = x[2 * q], r1 = x[2 * q + 1];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const unsigned int u = __builtin_bit_cast(unsigned int, __builtin_convertvector(f2v{r0, r1}, bf2v));
      w[i][q] = u;
      if (i + 1 < NS) {
        r0 -= __uint_as_float(u << 16);
        r1 -= __uint_as_float(u & 0xffff0000u);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i)
    *reinterpret_cast<u2v*>(bf_plane<NS>(base, op, c, i) + row * BKP + k4) = u2v{w[i][0], w[i][1]};
}

__device__ __forceinline__ bf8v bf_neg(bf8v v) {
  u4v u = __builtin_bit_cast(u4v, v);
  u ^= 0x80008000u;
  return __builtin_bit_cast(bf8v, u);
}

// One 32-deep stage of a wave's 32 x 32 tile: acc_re/acc_im (+)= sum of the split products of
// (xr + i sx xi)(yr + i yi) with sx = -1 (conj(x) y, CONJ) or +1 (x y); rows of x are the
// output rows (operand 0), of y the output columns (operand 1)
template <int NS, bool CONJ>
__device__ __forceinline__ void bf_stage(unsigned short* base, int wm, int wn, int li, int lq, f4v (&cr)[2][2],
                                         f4v (&ci)[2][2]) {
  bf8v xr[2][NS], xi[2][NS], yr[2][NS], yi[2][NS];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const int rx = (wm + 16 * t + li) * BKP + 8 * lq, ry = (wn + 16 * t + li) * BKP + 8 * lq;
      xr[t][i] = *reinterpret_cast<const bf8v*>(bf_plane<NS>(base, 0, 0, i) + rx);
      xi[t][i] = *reinterpret_cast<const bf8v*>(bf_plane<NS>(base, 0, 1, i) + rx);
      yr[t][i] = *reinterpret_cast<const bf8v*>(bf_plane<NS>(base, 1, 0, i) + ry);
      yi[t][i] = *reinterpret_cast<const bf8v*>(bf_plane<NS>(base, 1, 1, i) + ry);
    }
  // CONJ:  Re = xr yr + xi yi, Im = xr yi - xi yr;  else Re = xr yr - xi yi, Im = xr yi + xi yr
#pragma unroll
  for (int p = 0; p < NS; ++p)
#pragma unroll
    for (int q = 0; p + q < NS; ++q) {
      bf8v nx[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) nx[t] = bf_neg(xi[t][p]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cr[a][b] = mfma32bf(xr[a][p], yr[b][q], cr[a][b]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) ci[a][b] = mfma32bf(xr[a][p], yi[b][q], ci[a][b]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cr[a][b] = mfma32bf(CONJ ? xi[a][p] : nx[a], yi[b][q], cr[a][b]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) ci[a][b] = mfma32bf(CONJ ? nx[a] : xi[a][p], yr[b][q], ci[a][b]);
    }
}

// proj: part[s][m][n] = sum_{k in chunk s} conj(A[m][k]) B[n][k]  (same tiling, split-K
// decoding and XCD-aware order as k_gemm_proj)
template <int NS>
__global__ __launch_bounds__(GT) void k_gemm_proj_bf(const cx<float>* __restrict__ A, int lda,
                                                     const cx<float>* __restrict__ B, int ldb, int M, int N,
                                                     int K, int kchunk, cx<float>* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char gsm[];
  unsigned short* base = reinterpret_cast<unsigned short*>(gsm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tiles_n = (N + BT - 1) / BT, tiles = ((M + BT - 1) / BT) * tiles_n;
  const int L = blockIdx.x, q = L >> 3;
  const bool xcd = ((K + kchunk - 1) / kchunk) % 8 == 0;
  const int tile = xcd ? q % tiles : L % tiles;
  const int split = xcd ? (L & 7) + 8 * (q / tiles) : L / tiles;
  if (split * kchunk >= K) return;
  const int m0 = (tile / tiles_n) * BT, n0 = (tile % tiles_n) * BT;
  const int ks = split * kchunk, ke = min(K, ks + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  f4v cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) { cr[a][b] = f4v{0, 0, 0, 0}; ci[a][b] = f4v{0, 0, 0, 0}; }
  // staging: thread -> rows (t + 256 p) / 8 (p = 0, 1) of both operands, 4 consecutive k at
  // 4 (t % 8), read as two 16-byte loads (global loads need 4-byte alignment only, so any
  // lda works): a wave's load touches 16 rows of 32 bytes x 8 lanes.  Rows past the matrix
  // read row 0 and k past the chunk the chunk's last column, both then replaced by zero, so
  // every load is in bounds and the loop has no branches.
  const int sk = 4 * (tid & 7);
  int srow[2];
  bool arow[2], brow[2];
  const cx<float>* ap[2];
  const cx<float>* bp[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    srow[p] = (tid + GT * p) >> 3;
    arow[p] = m0 + srow[p] < M;
    brow[p] = n0 + srow[p] < N;
    ap[p] = A + (size_t)(arow[p] ? m0 + srow[p] : 0) * lda;
    bp[p] = B + (size_t)(brow[p] ? n0 + srow[p] : 0) * ldb;
  }
  f4v pa[2][2], pb[2][2];        // [p][h]: complex k0 + sk + 2h, + 1 as (re, im, re, im)
  auto fetch = [&](int k0) {
    if (k0 + BKS <= ke) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          f4v va, vb;
          __builtin_memcpy(&va, ap[p] + k0 + sk + 2 * h, 16);
          __builtin_memcpy(&vb, bp[p] + k0 + sk + 2 * h, 16);
          pa[p][h] = arow[p] ? va : f4v{0, 0, 0, 0};
          pb[p][h] = brow[p] ? vb : f4v{0, 0, 0, 0};
        }
    } else {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = k0 + sk + e, kc = min(k, ke - 1);
          const cx<float> va = ap[p][kc], vb = bp[p][kc];
          const bool ka = arow[p] && k < ke, kb = brow[p] && k < ke;
          pa[p][e >> 1][2 * (e & 1)] = ka ? va.x : 0.f;
          pa[p][e >> 1][2 * (e & 1) + 1] = ka ? va.y : 0.f;
          pb[p][e >> 1][2 * (e & 1)] = kb ? vb.x : 0.f;
          pb[p][e >> 1][2 * (e & 1) + 1] = kb ? vb.y : 0.f;
        }
    }
  };
  fetch(ks);
  for (int k0 = ks; k0 < ke; k0 += BKS) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const float ar[4] = {pa[p][0][0], pa[p][0][2], pa[p][1][0], pa[p][1][2]};
      const float ai[4] = {pa[p][0][1], pa[p][0][3], pa[p][1][1], pa[p][1][3]};
      const float br[4] = {pb[p][0][0], pb[p][0][2], pb[p][1][0], pb[p][1][2]};
      const float bi[4] = {pb[p][0][1], pb[p][0][3], pb[p][1][1], pb[p][1][3]};
      bf_store4<NS>(base, 0, 0, srow[p], sk, ar);
      bf_store4<NS>(base, 0, 1, srow[p], sk, ai);
      bf_store4<NS>(base, 1, 0, srow[p], sk, br);
      bf_store4<NS>(base, 1, 1, srow[p], sk, bi);
    }
    __syncthreads();
    if (k0 + BKS < ke) fetch(k0 + BKS);
    bf_stage<NS, true>(base, wm, wn, li, lq, cr, ci);
    __syncthreads();
  }
  cx<float>* P = part + (size_t)split * M * N;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + 4 * lq + r, n = n0 + wn + b * 16 + li;
        if (m < M && n < N) P[(size_t)m * N + n] = mk<float>(cr[a][b][r], ci[a][b][r]);
      }
}

// expand: out[n][k] (+)= sum_j A[j][k] Q[j][n]  (same tiling and order as k_gemm_expand);
// the staging transposes both tiles so that j is contiguous in LDS
template <int NS>
__global__ __launch_bounds__(GT) void k_gemm_expand_bf(const cx<float>* __restrict__ A, int lda,
                                                       const cx<float>* __restrict__ Q, int ldq, int J, int N,
                                                       int K, cx<float>* __restrict__ out, int ldo, int accumulate) {
  extern __shared__ __align__(16) unsigned char gsm[];
  unsigned short* base = reinterpret_cast<unsigned short*>(gsm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tiles_n = (N + BT - 1) / BT;
  const int L = blockIdx.x, q = L >> 3;
  const int k0 = ((L & 7) + 8 * (q / tiles_n)) * BT, n0 = (q % tiles_n) * BT;
  if (k0 >= K) return;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  f4v cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) { cr[a][b] = f4v{0, 0, 0, 0}; ci[a][b] = f4v{0, 0, 0, 0}; }
  // staging: thread -> column t % 64 (a plane wave of A, a band of Q), 8 consecutive j at
  // 8 (t / 64); consecutive threads read consecutive addresses of a row
  const int sc = tid & 63, sj = 8 * (tid >> 6);
  const bool kin = k0 + sc < K, nin = n0 + sc < N;
  cx<float> pa[8], pq[8];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int j = j0 + sj + e;
      pa[e] = (kin && j < J) ? A[(size_t)j * lda + k0 + sc] : mk<float>(0, 0);
      pq[e] = (nin && j < J) ? Q[(size_t)j * ldq + n0 + sc] : mk<float>(0, 0);
    }
  };
  fetch(0);
  for (int j0 = 0; j0 < J; j0 += BKS) {
    {
      float ar[8], ai[8], qr[8], qi[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) { ar[e] = pa[e].x; ai[e] = pa[e].y; qr[e] = pq[e].x; qi[e] = pq[e].y; }
      bf_store8<NS>(base, 0, 0, sc, sj, ar);
      bf_store8<NS>(base, 0, 1, sc, sj, ai);
      bf_store8<NS>(base, 1, 0, sc, sj, qr);
      bf_store8<NS>(base, 1, 1, sc, sj, qi);
    }
    __syncthreads();
    if (j0 + BKS < J) fetch(j0 + BKS);
    bf_stage<NS, false>(base, wm, wn, li, lq, cr, ci);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + wm + a * 16 + 4 * lq + r, n = n0 + wn + b * 16 + li;
        if (k < K && n < N) {
          cx<float> v = mk<float>(cr[a][b][r], ci[a][b][r]);
          cx<float>* o = out + (size_t)n * ldo + k;
          if (accumulate) v = v + *o;
          *o = v;
        }
      }
}

// ---------------------------------------------------------------------------
// One band (nb = 1, the single-vector H|psi> of the Lanczos bounds and ApplyHam): the
// products are matrix-vector, bound by reading beta once, so they skip the 64-wide MFMA
// tiles.  Both reduce in a fixed order (deterministic).
// ---------------------------------------------------------------------------
// C[m] = wl[m % lm] sum_k conj(A[m][k]) x[k], one block per row
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_proj(const cx<T>* __restrict__ A, int lda, const cx<T>* __restrict__ x,
                                                   int K, cx<T>* __restrict__ C, const T* __restrict__ wl, int lm) {
  __shared__ T sr[256], si[256];
  const int m = blockIdx.x, t = threadIdx.x;
  const cx<T>* a = A + (size_t)m * lda;
  T ar0 = 0, ai0 = 0, ar1 = 0, ai1 = 0;
  int k = t;
  for (; k + 256 < K; k += 512) {
    const cx<T> u0 = a[k], v0 = x[k], u1 = a[k + 256], v1 = x[k + 256];
    ar0 += u0.x * v0.x + u0.y * v0.y; ai0 += u0.x * v0.y - u0.y * v0.x;
    ar1 += u1.x * v1.x + u1.y * v1.y; ai1 += u1.x * v1.y - u1.y * v1.x;
  }
  if (k < K) {
    const cx<T> u0 = a[k], v0 = x[k];
    ar0 += u0.x * v0.x + u0.y * v0.y; ai0 += u0.x * v0.y - u0.y * v0.x;
  }
  sr[t] = ar0 + ar1;
  si[t] = ai0 + ai1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) { sr[t] += sr[t + w]; si[t] += si[t + w]; }
    __syncthreads();
  }
  if (t == 0) {
    const T sc = wl ? wl[m % lm] : T(1);
    C[m] = mk<T>(sc * sr[0], sc * si[0]);
  }
}

// out[k] (+)= sum_m A[m][k] q[m]: a block takes 64 plane waves, its four waves the rows
// m = c, c + 4, ... (c = wave), added in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_expand(const cx<T>* __restrict__ A, int lda, int M,
                                                     const cx<T>* __restrict__ q, int K, cx<T>* __restrict__ out,
                                                     int accumulate) {
  __shared__ cx<T> part[4][64];
  const int kk = threadIdx.x & 63, c = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + kk;
  cx<T> acc = mk<T>(0, 0);
  if (k < K) {
#pragma unroll 8
    for (int m = c; m < M; m += 4) acc = acc + A[(size_t)m * lda + k] * q[m];
  }
  part[c][kk] = acc;
  __syncthreads();
  if (c == 0 && k < K) {
    cx<T> v = (part[0][kk] + part[1][kk]) + (part[2][kk] + part[3][kk]);
    if (accumulate) v = v + out[k];
    out[k] = v;
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int gemm_target_blocks() {
  static const int t = [] {
    const char* e = getenv("PWH_GEMM_BLOCKS");
    const int v = e ? atoi(e) : 0;
    return v >= 64 ? v : 512;
  }();
  return t;
}

static int choose_split(int tiles, int K, size_t part_elems, size_t MN, int target = 0) {
  // aim for ~2 blocks per CU on 256 CUs (or the tuned target), at least 128 plane waves
  // per block; a multiple of 8 where K allows, so the XCD-aware order keeps all 8 XCDs busy
  int s = ((target > 0 ? target : gemm_target_blocks()) + tiles - 1) / tiles;
  int smax = (K + 127) / 128;
  s = (s + 7) / 8 * 8;
  if (s > smax) s = smax;
  if (s < 1) s = 1;
  while (s > 1 && (size_t)s * MN > part_elems) --s;
  return s;
}

size_t proj_part_elems(int M, int N, int K, int target) {
  const int tiles = ((M + BT - 1) / BT) * ((N + BT - 1) / BT);
  const int s = choose_split(tiles, K, (size_t)-1, (size_t)M * N, target);
  return (size_t)s * M * N;
}

// Reduction depth per LDS stage: deeper stages give each wave more MFMA work per barrier
// and per register prefetch, which has to hide the global load latency.  PWH_GEMM_KS
// (16, 32, 64; 64 for FP32 only) overrides the default.
template <typename T>
static int gemm_ks() {
  static const int ks = [] {
    const char* e = getenv("PWH_GEMM_KS");
    const int v = e ? atoi(e) : 0;
    if (v == 16 || v == 32 || (v == 64 && sizeof(T) == 4)) return v;
    return 16;
  }();
  return ks;
}

// dynamic LDS above 64 KiB has to be enabled once per kernel
template <typename F>
static hipError_t allow_smem(F* kern, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

// 1D grid of k_gemm_expand: plane-wave tiles rounded up to 8, times the band tiles
static int expand_blocks(int K, int N) {
  const int tk = (K + BT - 1) / BT, tn = (N + BT - 1) / BT;
  return ((tk + 7) / 8) * 8 * tn;
}

template <typename T, bool BETA, int KS>
static hipError_t launch_proj_ks(dim3 grid, hipStream_t s, RowLoader<T, BETA> ld, const cx<T>* B, int ldb,
                                 int M, int N, int K, int kchunk, cx<T>* part) {
  constexpr size_t sm = gemm_smem<T, KS>();
  static hipError_t attr = allow_smem(&k_gemm_proj<T, BETA, KS>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_proj<T, BETA, KS>), grid, dim3(GT), sm, s, ld, B, ldb, M, N, K, kchunk, part);
  return hipGetLastError();
}

template <typename T, bool BETA>
static hipError_t launch_proj(int ks, dim3 grid, hipStream_t s, RowLoader<T, BETA> ld, const cx<T>* B,
                              int ldb, int M, int N, int K, int kchunk, cx<T>* part) {
  if (ks == 32) return launch_proj_ks<T, BETA, 32>(grid, s, ld, B, ldb, M, N, K, kchunk, part);
  if constexpr (sizeof(T) == 4)
    if (ks == 64) return launch_proj_ks<T, BETA, 64>(grid, s, ld, B, ldb, M, N, K, kchunk, part);
  return launch_proj_ks<T, BETA, 16>(grid, s, ld, B, ldb, M, N, K, kchunk, part);
}

template <typename T, bool BETA, bool QCM, int KS>
static hipError_t launch_expand_ks(dim3 grid, hipStream_t s, RowLoader<T, BETA> ld, const cx<T>* Q, int ldq,
                                   int J, int N, int K, cx<T>* out, int ldo, int acc) {
  constexpr size_t sm = gemm_smem<T, KS>();
  static hipError_t attr = allow_smem(&k_gemm_expand<T, BETA, QCM, KS>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_expand<T, BETA, QCM, KS>), grid, dim3(GT), sm, s, ld, Q, ldq, J, N, K, out,
                     ldo, acc);
  return hipGetLastError();
}

template <typename T, bool BETA, bool QCM>
static hipError_t launch_expand(int ks, dim3 grid, hipStream_t s, RowLoader<T, BETA> ld, const cx<T>* Q,
                                int ldq, int J, int N, int K, cx<T>* out, int ldo, int acc) {
  if (ks == 32) return launch_expand_ks<T, BETA, QCM, 32>(grid, s, ld, Q, ldq, J, N, K, out, ldo, acc);
  if constexpr (sizeof(T) == 4)
    if (ks == 64) return launch_expand_ks<T, BETA, QCM, 64>(grid, s, ld, Q, ldq, J, N, K, out, ldo, acc);
  return launch_expand_ks<T, BETA, QCM, 16>(grid, s, ld, Q, ldq, J, N, K, out, ldo, acc);
}

// terms of the bf16 split of a precision mode (0: not a split mode)
static int bf_terms(int mode) { return mode == PREC_BF16X3 ? 2 : mode == PREC_BF16X6 ? 3 : 0; }

template <int NS>
static hipError_t launch_proj_bf(dim3 grid, hipStream_t s, const cx<float>* A, int lda, const cx<float>* B, int ldb,
                                 int M, int N, int K, int kchunk, cx<float>* part) {
  constexpr size_t sm = bf_smem<NS>();
  static hipError_t attr = allow_smem(&k_gemm_proj_bf<NS>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_proj_bf<NS>), grid, dim3(GT), sm, s, A, lda, B, ldb, M, N, K, kchunk, part);
  return hipGetLastError();
}

template <int NS>
static hipError_t launch_expand_bf(dim3 grid, hipStream_t s, const cx<float>* A, int lda, const cx<float>* Q, int ldq,
                                   int J, int N, int K, cx<float>* out, int ldo, int acc) {
  constexpr size_t sm = bf_smem<NS>();
  static hipError_t attr = allow_smem(&k_gemm_expand_bf<NS>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_expand_bf<NS>), grid, dim3(GT), sm, s, A, lda, Q, ldq, J, N, K, out, ldo, acc);
  return hipGetLastError();
}

// Split-K shape: S chunks of kchunk plane waves (a multiple of the stage depth ks), and
// the 1D grid (tiles times S rounded up to 8, decoded XCD-aware in the kernel)
static void split_shape(int M, int N, int K, size_t part_elems, int target, int ks, int* S_out, int* kchunk_out,
                        int* blocks) {
  const int tiles = ((M + BT - 1) / BT) * ((N + BT - 1) / BT);
  int S = choose_split(tiles, K, part_elems, (size_t)M * N, target);
  int kchunk = (K + S - 1) / S;
  kchunk = ((kchunk + ks - 1) / ks) * ks;
  S = (K + kchunk - 1) / kchunk;
  *S_out = S;
  *kchunk_out = kchunk;
  *blocks = tiles * ((S + 7) / 8) * 8;
}

template <typename T>
int proj_split_shape(int M, int N, int K, int target, int* S, int* kchunk, int* blocks) {
  if (gemm_ks<T>() != 16) return 0;
  split_shape(M, N, K, proj_part_elems(M, N, K, target), target, 16, S, kchunk, blocks);
  return 1;
}

int expand_grid_blocks(int K, int N) { return expand_blocks(K, N); }

template <typename T>
hipError_t launch_group_proj(const GroupProj<T>* g, int n, int blocks, int rblocks, hipStream_t s) {
  if (n < 1 || blocks < 1 || rblocks < 1) return hipErrorInvalidValue;
  constexpr size_t sm = gemm_smem<T, 16>();
  static hipError_t attr = allow_smem(&k_gemm_proj_group<T, 16>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_proj_group<T, 16>), dim3(blocks), dim3(GT), sm, s, g, n);
  hipLaunchKernelGGL((k_reduce_group<T>), dim3(rblocks), dim3(256), 0, s, g, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_group_expand(const GroupExpand<T>* g, int n, int blocks, hipStream_t s) {
  if (n < 1 || blocks < 1) return hipErrorInvalidValue;
  constexpr size_t sm = gemm_smem<T, 16>();
  static hipError_t attr = allow_smem(&k_gemm_expand_group<T, 16>, sm);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_gemm_expand_group<T, 16>), dim3(blocks), dim3(GT), sm, s, g, n);
  return hipGetLastError();
}

template int proj_split_shape<double>(int, int, int, int, int*, int*, int*);
template int proj_split_shape<float>(int, int, int, int, int*, int*, int*);
template hipError_t launch_group_proj<double>(const GroupProj<double>*, int, int, int, hipStream_t);
template hipError_t launch_group_proj<float>(const GroupProj<float>*, int, int, int, hipStream_t);
template hipError_t launch_group_expand<double>(const GroupExpand<double>*, int, int, hipStream_t);
template hipError_t launch_group_expand<float>(const GroupExpand<float>*, int, int, hipStream_t);

// split-K projection: S chunks of kchunk plane waves (a multiple of the stage depth)
template <typename T, bool BETA>
static hipError_t run_proj(RowLoader<T, BETA> ld, const cx<T>* B, int ldb, int M, int N, int K,
                           cx<T>* part, size_t part_elems, int mode, hipStream_t s, int* S_out,
                           int target = 0, int ks_set = 0) {
  const int nsplit = std::is_same<T, float>::value ? bf_terms(mode) : 0;
  const int ks = nsplit ? BKS : (ks_set ? ks_set : gemm_ks<T>());
  int S, kchunk, blocks;
  split_shape(M, N, K, part_elems, target, ks, &S, &kchunk, &blocks);
  if ((size_t)S * M * N > part_elems) return hipErrorInvalidValue;
  dim3 grid(blocks);   // 1D, decoded XCD-aware in the kernel
  hipError_t e;
  if constexpr (std::is_same<T, float>::value) {
    if (nsplit == 3) e = launch_proj_bf<3>(grid, s, ld.A, ld.lda, B, ldb, M, N, K, kchunk, part);
    else if (nsplit == 2) e = launch_proj_bf<2>(grid, s, ld.A, ld.lda, B, ldb, M, N, K, kchunk, part);
    else e = launch_proj<T, BETA>(ks, grid, s, ld, B, ldb, M, N, K, kchunk, part);
  } else {
    e = launch_proj<T, BETA>(ks, grid, s, ld, B, ldb, M, N, K, kchunk, part);
  }
  *S_out = S;
  return e;
}

template <typename T>
hipError_t gemm_proj(const BetaGen& bg, const cx<T>* X, int ldx, int nb, int K, cx<T>* C,
                     cx<T>* part, size_t part_elems, int mode, hipStream_t s, const T* wl) {
  RowLoader<T, false> ld{};
  ld.A = static_cast<const cx<T>*>(bg.B);
  ld.lda = bg.ld;
  const int M = bg.rows;
  if (nb == 1) {
    hipLaunchKernelGGL((k_gemv_proj<T>), dim3(M), dim3(256), 0, s, ld.A, ld.lda, X, K, C, wl, bg.lm);
    return hipGetLastError();
  }
  int S = 0;
  hipError_t e;
  if constexpr (std::is_same<T, float>::value) {
    if (bg.fly && mode == PREC_FP32) {
      RowLoader<float, true> lf{nullptr, 0, static_cast<const cx<float>*>(bg.ei), bg.mi, bg.pkg, bg.ldt, bg.lm,
                                bg.ld, bg.slot0};
      e = run_proj<float, true>(lf, X, ldx, M, nb, K, part, part_elems, mode, s, &S, bg.target, bg.ks);
    } else {
      e = run_proj<T, false>(ld, X, ldx, M, nb, K, part, part_elems, mode, s, &S, bg.target, bg.ks);
    }
  } else {
    e = run_proj<T, false>(ld, X, ldx, M, nb, K, part, part_elems, mode, s, &S, bg.target, bg.ks);
  }
  if (e != hipSuccess) return e;
  const int tot = M * nb;
  hipLaunchKernelGGL((k_reduce_parts<T, T>), dim3((tot + 255) / 256), dim3(256), 0, s, part, S, M,
                     nb, C, nb, wl, bg.lm, 0);
  return hipGetLastError();
}

template <typename T>
hipError_t gemm_expand(const BetaGen& bg, const cx<T>* Q, int nb, int K, cx<T>* out, int ldo,
                       bool accumulate, int mode, hipStream_t s) {
  RowLoader<T, false> ld{};
  ld.A = static_cast<const cx<T>*>(bg.B);
  ld.lda = bg.ld;
  const int J = bg.rows;
  if (nb == 1) {
    hipLaunchKernelGGL((k_gemv_expand<T>), dim3((K + 63) / 64), dim3(256), 0, s, ld.A, ld.lda, J, Q, K, out,
                       accumulate ? 1 : 0);
    return hipGetLastError();
  }
  dim3 grid(expand_blocks(K, nb));
  const int acc = accumulate ? 1 : 0;
  if constexpr (std::is_same<T, float>::value) {
    const int nsplit = bf_terms(mode);
    if (nsplit == 3) return launch_expand_bf<3>(grid, s, ld.A, ld.lda, Q, nb, J, nb, K, out, ldo, acc);
    if (nsplit == 2) return launch_expand_bf<2>(grid, s, ld.A, ld.lda, Q, nb, J, nb, K, out, ldo, acc);
    if (bg.fly && mode == PREC_FP32) {
      RowLoader<float, true> lf{nullptr, 0, static_cast<const cx<float>*>(bg.ei), bg.mi, bg.pkg, bg.ldt, bg.lm,
                                bg.ld, bg.slot0};
      return launch_expand<float, true, false>(bg.ks ? bg.ks : gemm_ks<float>(), grid, s, lf, Q, nb, J, nb, K, out,
                                               ldo, acc);
    }
  }
  return launch_expand<T, false, false>(bg.ks ? bg.ks : gemm_ks<T>(), grid, s, ld, Q, nb, J, nb, K, out, ldo, acc);
}

hipError_t gemm_cn_f64(const cx<double>* A, int lda, int M, const cx<double>* B, int ldb, int N,
                       int K, cx<double>* C, int ldc, cx<double>* part, size_t part_elems,
                       hipStream_t s) {
  RowLoader<double, false> ld{};
  ld.A = A; ld.lda = lda;
  int S = 0;
  hipError_t e = run_proj<double, false>(ld, B, ldb, M, N, K, part, part_elems, PREC_FP64, s, &S);
  if (e != hipSuccess) return e;
  const int tot = M * N;
  hipLaunchKernelGGL((k_reduce_parts<double, double>), dim3((tot + 255) / 256), dim3(256), 0, s, part,
                     S, M, N, C, ldc, static_cast<const double*>(nullptr), 1, 1);
  return hipGetLastError();
}

// Split of the upper-triangle Gram product: the split count comes from the tiles that
// compute (the upper ones), not from the full square of tiles.
static void gram_split_shape(int N, int K, size_t part_elems, int* S_out, int* kchunk_out, int* blocks) {
  const int tn = (N + BT - 1) / BT, tiles = tn * (tn + 1) / 2;
  int S = choose_split(tiles, K, part_elems, (size_t)N * N, 0);
  int kchunk = (K + S - 1) / S;
  kchunk = ((kchunk + 15) / 16) * 16;
  S = (K + kchunk - 1) / kchunk;
  *S_out = S;
  *kchunk_out = kchunk;
  *blocks = tiles * ((S + 7) / 8) * 8;
}

size_t gram_part_elems(int N, int K) {
  int S, kchunk, blocks;
  gram_split_shape(N, K, (size_t)-1, &S, &kchunk, &blocks);
  return (size_t)S * N * N;
}

hipError_t gemm_gram_upper_f64(const cx<double>* A, int lda, int N, int K, cx<double>* C, cx<double>* part,
                               size_t part_elems, hipStream_t s) {
  int S, kchunk, blocks;
  gram_split_shape(N, K, part_elems, &S, &kchunk, &blocks);
  if ((size_t)S * N * N > part_elems) return hipErrorInvalidValue;
  RowLoader<double, false> ld{A, lda};
  constexpr size_t sm = gemm_smem<double, 16>();
  hipLaunchKernelGGL((k_gemm_gram_upper<16>), dim3(blocks), dim3(GT), sm, s, ld, A, lda, N, N, K, kchunk, part);
  const int tot = N * N;
  hipLaunchKernelGGL((k_reduce_parts<double, double>), dim3((tot + 255) / 256), dim3(256), 0, s, part, S, N, N, C,
                     N, static_cast<const double*>(nullptr), 1, 1);
  return hipGetLastError();
}

hipError_t gemm_nn_tri_f64(const cx<double>* A, int lda, int J, const cx<double>* Q, int ldq, int N, int K,
                           cx<double>* out, int ldo, hipStream_t s) {
  RowLoader<double, false> ld{A, lda};
  constexpr size_t sm = gemm_smem<double, 16>();
  hipLaunchKernelGGL((k_gemm_nn_tri<16>), dim3(expand_blocks(K, N)), dim3(GT), sm, s, ld, Q, ldq, J, N, K, out,
                     ldo);
  return hipGetLastError();
}

hipError_t gemm_nn_f64(const cx<double>* A, int lda, int J, const cx<double>* Q, int ldq, int N,
                       int K, cx<double>* out, int ldo, hipStream_t s) {
  RowLoader<double, false> ld{};
  ld.A = A; ld.lda = lda;
  dim3 grid(expand_blocks(K, N));
  return launch_expand<double, false, true>(gemm_ks<double>(), grid, s, ld, Q, ldq, J, N, K, out, ldo, 0);
}

// ---------------------------------------------------------------------------
// Self-test of the MFMA operand / result lane layouts assumed above: one wave
// computes D = A(16x4) B(4x16) (kind 0: f64, 1: f32) or A(16x32) B(32x16)
// (kind 2: bf16) and the host compares with a plain product.
// ---------------------------------------------------------------------------
__global__ void k_mfma_probe(int kind, const double* A, const double* B, double* D) {
  const int lane = threadIdx.x, li = lane & 15, lq = lane >> 4;
  if (kind == 0) {
    d4v c = {0, 0, 0, 0};
    c = mfma4(A[li * 4 + lq], B[lq * 16 + li], c);
    for (int r = 0; r < 4; ++r) D[acc_row<double>(lq, r) * 16 + li] = c[r];
  } else if (kind == 1) {
    f4v c = {0, 0, 0, 0};
    c = mfma4((float)A[li * 4 + lq], (float)B[lq * 16 + li], c);
    for (int r = 0; r < 4; ++r) D[acc_row<float>(lq, r) * 16 + li] = c[r];
  } else {
    // 16x16x32 bf16: lane (li, lq) holds A[li][8 lq + v] and B[8 lq + v][li], v < 8
    f4v c = {0, 0, 0, 0};
    bf8v a, b;
    for (int v = 0; v < 8; ++v) {
      a[v] = __builtin_bit_cast(__bf16, f2bf((float)A[li * 32 + 8 * lq + v]));
      b[v] = __builtin_bit_cast(__bf16, f2bf((float)B[(8 * lq + v) * 16 + li]));
    }
    c = mfma32bf(a, b, c);
    for (int r = 0; r < 4; ++r) D[acc_row<float>(lq, r) * 16 + li] = c[r];
  }
}

hipError_t mfma_selftest(int kind, double* max_err) {
  const int KD = (kind == 2) ? 32 : 4;
  std::vector<double> hA(16 * KD), hB(KD * 16), hD(256), ref(256, 0.0);
  for (int i = 0; i < 16 * KD; ++i) hA[i] = (double)((i * 7) % 13) - 6.0;   // small integers: exact
  for (int i = 0; i < KD * 16; ++i) hB[i] = (double)((i * 5) % 11) - 5.0;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j)
      for (int k = 0; k < KD; ++k) ref[i * 16 + j] += hA[i * KD + k] * hB[k * 16 + j];
  double *dA = nullptr, *dB = nullptr, *dD = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&dA, sizeof(double) * hA.size())) != hipSuccess) return e;
  if ((e = hipMalloc(&dB, sizeof(double) * hB.size())) != hipSuccess) { hipFree(dA); return e; }
  if ((e = hipMalloc(&dD, sizeof(double) * 256)) != hipSuccess) { hipFree(dA); hipFree(dB); return e; }
  hipMemcpy(dA, hA.data(), sizeof(double) * hA.size(), hipMemcpyHostToDevice);
  hipMemcpy(dB, hB.data(), sizeof(double) * hB.size(), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, kind, dA, dB, dD);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(hD.data(), dD, sizeof(double) * 256, hipMemcpyDeviceToHost);
  hipFree(dA); hipFree(dB); hipFree(dD);
  double m = 0;
  for (int i = 0; i < 256; ++i) m = fmax(m, fabs(hD[i] - ref[i]));
  *max_err = m;
  return e;
}

template hipError_t gemm_proj<double>(const BetaGen&, const cx<double>*, int, int, int, cx<double>*,
                                      cx<double>*, size_t, int, hipStream_t, const double*);
template hipError_t gemm_proj<float>(const BetaGen&, const cx<float>*, int, int, int, cx<float>*,
                                     cx<float>*, size_t, int, hipStream_t, const float*);
template hipError_t gemm_expand<double>(const BetaGen&, const cx<double>*, int, int, cx<double>*, int,
                                        bool, int, hipStream_t);
template hipError_t gemm_expand<float>(const BetaGen&, const cx<float>*, int, int, cx<float>*, int,
                                       bool, int, hipStream_t);

}  // namespace pwh
