// pwh_chol.hip -- inverse Cholesky factor of a small Hermitian positive definite matrix
// in one workgroup: S = R^H R (upper R, positive real diagonal) is replaced by R^{-1}
// (upper, zeros below).  This is the zpotrf('U') + ztrtri('U', 'N') pair of CholeskyQR
// (ortho_qr, ChebySCFLoop.cpp's orthonormalisation step) and of the block solvers'
// Cholesky orthonormalisation (lobpcg_cpu.cpp:126-166), for n <= 128: one launch in
// place of rocSOLVER's many small ones, and no host check between the two.
//
// Blocked by 4 x 4, with the matrix in registers: wave w takes the 8 x 8 blocks of tile
// (w / 4, w % 4), lane l of it the block (BI, BJ) = (8 (w / 4) + l / 8, 8 (w % 4) + l % 8),
// i.e. the entries (4 BI + r, 4 BJ + c), r, c < 4; the six tiles below the diagonal idle.
// Beyond n the matrix is padded with the identity, so every block is full.
//  * Factorisation, right-looking by block steps K (three barriers each): the lane of the
//    diagonal block factors it and inverts its triangle in registers (V = R_KK^{-1}); the
//    lanes of block row K form R_KJ = V^H A_KJ; every lane below updates
//    A_IJ -= R_KI^H R_KJ from the block row published in LDS.  R is kept packed by blocks
//    in LDS.  A pivot that is not positive (or not finite) stops the factorisation with
//    info = index + 1, as zpotrf does.
//  * Inverse, right-looking on B = I by block steps I = nb-1..0: the lanes of block row I
//    form U_IJ = V_II B_IJ and publish it, every lane above updates B_LJ -= R_LI U_IJ.
#include "pwh_internal.h"

namespace pwh {

constexpr int kCholThreads = 1024;
constexpr int kCholBlocks = kCholSmallMax / 4;   // 4 x 4 blocks per dimension

// 1/sqrt(a) for a > 0: the hardware estimate refined by two Newton steps (quadratic
// convergence from ~23 bits)
__device__ __forceinline__ double rsqrt_refined(double a) {
  double y = __builtin_amdgcn_rsq(a);
  y = y * fma(-0.5 * a * y, y, 1.5);
  y = y * fma(-0.5 * a * y, y, 1.5);
  return y;
}

__device__ __forceinline__ int bpk(int I, int J) { return J * (J + 1) / 2 + I; }   // I <= J

// LDS: R blocks (packed upper), the diagonal-block inverses, a double-buffered block row,
// the pivot flag
size_t chol_small_smem() {
  return sizeof(double2) * 16 * ((size_t)kCholBlocks * (kCholBlocks + 1) / 2 + 3 * kCholBlocks) + 16;
}

__global__ __launch_bounds__(kCholThreads) void k_chol_inv(cx<double>* __restrict__ S, int lda, int n,
                                                           int* __restrict__ info) {
  extern __shared__ __align__(16) unsigned char csm[];
  double2* Rb = reinterpret_cast<double2*>(csm);                 // R blocks, packed upper, [blk][4][4]
  double2* dinv = Rb + 16 * (kCholBlocks * (kCholBlocks + 1) / 2);   // R_KK^{-1} per block
  double2* prow = dinv + 16 * kCholBlocks;                       // [2][blocks][4][4]
  int* flag = reinterpret_cast<int*>(prow + 32 * kCholBlocks);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int tr = w >> 2, tc = w & 3;
  const int BI = 8 * tr + (lane >> 3), BJ = 8 * tc + (lane & 7);
  const int nbk = (n + 3) >> 2;
  const bool live = BI <= BJ && BJ < nbk;          // upper block in range
  const bool wave_live = tr <= tc && 8 * tc < nbk;
  double ar[4][4], ai[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * BI + r, j = 4 * BJ + c;
      const cx<double> v = S[min(i, n - 1) + (size_t)min(j, n - 1) * lda];
      const bool in = live && i < n && j < n;
      // identity padding beyond n
      ar[r][c] = in ? v.x : ((live && i == j) ? 1.0 : 0.0);
      ai[r][c] = in ? v.y : 0.0;
    }
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int K = 0; K < nbk; ++K) {
    double2* pr = prow + (K & 1) * 16 * kCholBlocks;
    // 1. diagonal block: local Cholesky and inverse of R_KK
    if (live && BI == K && BJ == K) {
      int bad = 0;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const double app = ar[p][p];
        if (!bad && (!(app > 0.0) || !(app < INFINITY))) bad = 4 * K + p + 1;
        const double rd = bad ? 1.0 : rsqrt_refined(app);
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c > p) { ar[p][c] *= rd; ai[p][c] *= rd; }
        ar[p][p] = bad ? 1.0 : app * rd; ai[p][p] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if (r > p && c >= r) {   // a_rc -= conj(a_pr) a_pc
              ar[r][c] -= ar[p][r] * ar[p][c] + ai[p][r] * ai[p][c];
              ai[r][c] -= ar[p][r] * ai[p][c] - ai[p][r] * ar[p][c];
            }
      }
      // lower part of the block is not part of R
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (r > c) { ar[r][c] = 0.0; ai[r][c] = 0.0; }
      // V = R_KK^{-1} (upper), one column at a time by back substitution, straight to LDS
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double vr[4], vi[4];
#pragma unroll
        for (int r = 3; r >= 0; --r) {
          double sr = (r == c) ? 1.0 : 0.0, si = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q > r && q <= c) {
              sr -= ar[r][q] * vr[q] - ai[r][q] * vi[q];
              si -= ar[r][q] * vi[q] + ai[r][q] * vr[q];
            }
          const double inv = 1.0 / ar[r][r];
          vr[r] = (r <= c) ? sr * inv : 0.0;
          vi[r] = (r <= c) ? si * inv : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) dinv[16 * K + 4 * r + c] = make_double2(vr[r], vi[r]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          pr[16 * K + 4 * r + c] = make_double2(ar[r][c], ai[r][c]);
          Rb[16 * bpk(K, K) + 4 * r + c] = make_double2(ar[r][c], ai[r][c]);
        }
      if (bad) *flag = bad;
    }
    __syncthreads();
    if (*flag) break;
    // 2. panel: R_KJ = R_KK^{-H} A_KJ, out[r][c] = sum_p conj(V[p][r]) A[p][c]
    if (live && BI == K && BJ > K) {
      // in place from the last row up: row r needs rows p <= r (V upper)
#pragma unroll
      for (int r = 3; r >= 0; --r) {
        double2 v[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) v[p] = dinv[16 * K + 4 * p + r];   // V[p][r]
        double orr[4], oii[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          orr[c] = 0.0; oii[c] = 0.0;
#pragma unroll
          for (int p = 0; p < 4; ++p)
            if (p <= r) {
              orr[c] = fma(v[p].x, ar[p][c], fma(v[p].y, ai[p][c], orr[c]));
              oii[c] = fma(v[p].x, ai[p][c], fma(-v[p].y, ar[p][c], oii[c]));
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          ar[r][c] = orr[c]; ai[r][c] = oii[c];
          pr[16 * BJ + 4 * r + c] = make_double2(orr[c], oii[c]);
          Rb[16 * bpk(K, BJ) + 4 * r + c] = make_double2(orr[c], oii[c]);
        }
      }
    }
    __syncthreads();
    // 3. trailing update: A_IJ -= R_KI^H R_KJ for K < I <= J
    if (wave_live && 8 * tr + 7 > K && live && BI > K) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        double2 x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          x[q] = pr[16 * BI + 4 * p + q];
          y[q] = pr[16 * BJ + 4 * p + q];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {   // a -= conj(x_r) y_c, as four FMAs
            ar[r][c] = fma(-x[r].y, y[c].y, fma(-x[r].x, y[c].x, ar[r][c]));
            ai[r][c] = fma(x[r].y, y[c].x, fma(-x[r].x, y[c].y, ai[r][c]));
          }
      }
    }
  }
  if (*flag) {
    if (tid == 0) info[0] = *flag;
    return;
  }
  // inverse: B = I; for I = nbk-1..0: U_IJ = V_II B_IJ (row I owners, published),
  // then B_LJ -= R_LI U_IJ for L < I <= J
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) { ar[r][c] = (live && BI == BJ && r == c) ? 1.0 : 0.0; ai[r][c] = 0.0; }
  for (int I = nbk - 1; I >= 0; --I) {
    double2* ub = prow + (I & 1) * 16 * kCholBlocks;
    if (live && BI == I) {
      // U = V B in place from the first row down: row r needs rows q >= r (V upper)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double2 v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = dinv[16 * I + 4 * r + q];   // V[r][q]
        double orr[4], oii[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          orr[c] = 0.0; oii[c] = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q >= r) {
              orr[c] = fma(v[q].x, ar[q][c], fma(-v[q].y, ai[q][c], orr[c]));
              oii[c] = fma(v[q].x, ai[q][c], fma(v[q].y, ar[q][c], oii[c]));
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          ar[r][c] = orr[c]; ai[r][c] = oii[c];
          ub[16 * BJ + 4 * r + c] = make_double2(orr[c], oii[c]);
        }
      }
    }
    __syncthreads();
    if (wave_live && 8 * tr < I && live && BI < I && BJ >= I) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double2 x[4], y[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          x[t] = Rb[16 * bpk(BI, I) + 4 * t + q];   // R_LI[r=t][q]
          y[t] = ub[16 * BJ + 4 * q + t];           // U_IJ[q][c=t]
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {   // b -= x_r y_c, as four FMAs
            ar[r][c] = fma(x[r].y, y[c].y, fma(-x[r].x, y[c].x, ar[r][c]));
            ai[r][c] = fma(-x[r].y, y[c].x, fma(-x[r].x, y[c].y, ai[r][c]));
          }
      }
    }
  }
  if (live) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = 4 * BI + r, j = 4 * BJ + c;
        if (i < n && j < n) S[i + (size_t)j * lda] = (i <= j) ? mk<double>(ar[r][c], ai[r][c]) : mk<double>(0.0, 0.0);
      }
  }
  for (int e = tid; e < n * n; e += kCholThreads) {
    const int i = e % n, j = e / n;
    if ((i >> 2) > (j >> 2)) S[i + (size_t)j * lda] = mk<double>(0.0, 0.0);
  }
  if (tid == 0) info[0] = 0;
}

hipError_t launch_chol_inv_small(cx<double>* S, int lda, int n, int* info, hipStream_t s) {
  if (n < 1 || n > kCholSmallMax || lda < n) return hipErrorInvalidValue;
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_chol_inv),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)chol_small_smem());
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(kCholThreads), chol_small_smem(), s, S, lda, n, info);
  return hipGetLastError();
}

}  // namespace pwh
