// pwh_fft_rt.hip -- the FFT passes for any grid: runtime-length Stockham plans.
//
// The reference plans any NR1 x NR2 x NR3 through FFTW (fft_fftw3.cpp:29,35), and its
// Setup picks the smallest even size above the cutoff minimum with no smoothing
// (Setup.cpp:20-34), so grids like 22 = 2*11, 28 = 4*7 or 70 x 70 x 140 occur in its own
// runs.  The compiled per-length kernels (pwh_fft.hip) cover 2,3,5-smooth lengths up to
// 128 and square planes; everything else runs here:
//
//   * a plan is a list of radices (fft_plan); stages of radix 2, 3, 4, 5, 7, 8 hold their
//     butterflies in registers across one barrier (as the compiled kernels do), a stage of
//     any other prime p computes each output as a length-p sum read from LDS;
//   * k_zinv_rt / k_zfwd_rt: the i3 transforms of the sphere columns (gather, inverse;
//     forward, epilogue), CT columns per block;
//   * k_plane_rt: one (i3, band) plane per workgroup in LDS, any n1 x n2;
//   * split plane pass, for planes too large for LDS: the i2 transforms of the active i1
//     columns into an intermediate Y[band][i3][active i1][i2] in HBM (k_ysplit_rt), the
//     i1 rows with V between (k_xsplit_rt), and back (k_ysplit_rt), in chunks of bands;
//   * the density pass (k_plane_rho_rt, or k_ysplit_rt + k_xrho_split) and dense lines
//     (k_lines_rt) for pwh_fft3d and the density grid.
#include <utility>

#include "pwh_dft.h"

namespace pwh {

constexpr int kRtEmax = 16;   // transform elements a lane holds across a stage's barrier

bool fft_plan(int n, FftPlan* p) {
  if (!p || n < 2 || n > kFftMaxLen) return false;
  p->n = n;
  p->nst = 0;
  int rem = n;
  while (rem > 1) {
    int r;
    if (rem % 8 == 0 && rem != 16) r = 8;
    else if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else if (rem % 5 == 0) r = 5;
    else if (rem % 7 == 0) r = 7;
    else {
      r = 11;                      // rem has no factor 2, 3, 5 or 7 left
      while (rem % r != 0) r += 2;
    }
    if (p->nst >= kFftMaxStages) return false;
    p->radix[p->nst++] = r;
    rem /= r;
  }
  return true;
}

bool fft_plan_variant(int n, int variant, FftPlan* p) {
  if (variant == 0 || variant == 2) {
    if (!fft_plan(n, p)) return false;
    if (variant == 2)
      for (int i = 0, j = p->nst - 1; i < j; ++i, --j) std::swap(p->radix[i], p->radix[j]);
    return true;
  }
  if (variant != 1 || !p || n < 2 || n > kFftMaxLen) return false;
  p->n = n;
  p->nst = 0;
  int rem = n;
  while (rem > 1) {
    int r;
    if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else if (rem % 5 == 0) r = 5;
    else if (rem % 7 == 0) r = 7;
    else {
      r = 11;
      while (rem % r != 0) r += 2;
    }
    if (p->nst >= kFftMaxStages) return false;
    p->radix[p->nst++] = r;
    rem /= r;
  }
  return true;
}

static bool reg_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8; }

// Every stage of the plan fits `count` transforms on nt lanes holding kRtEmax elements.
static bool rt_fits(const FftPlan& p, int count, int nt) {
  for (int st = 0; st < p.nst; ++st) {
    const int r = p.radix[st];
    if (reg_radix(r)) {
      const long long bfly = (long long)count * (p.n / r);
      if ((bfly + nt - 1) / nt > kRtEmax / r) return false;
    } else {
      if (((long long)count * p.n + nt - 1) / nt > kRtEmax) return false;
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Stages.  Element e of transform t lives at buf[base(t) + e * es]; consecutive lanes take
// consecutive transforms.  NS = product of the earlier radices.  Hooks as in pwh_dft.h:
// live_in on the first stage's inputs, live_out and post on the last stage's outputs.
// ---------------------------------------------------------------------------
template <typename T, int R, int SIGN, int NT, typename BaseF, typename Hooks>
__device__ __forceinline__ void rt_stage(cx<T>* buf, int N, int NS, bool first, bool last, int count, int es,
                                         BaseF base, const cx<T>* tw, int tid, const Hooks& hk) {
  constexpr int QM = kRtEmax / R;
  const int NB = N / R;
  const int total = count * NB;
  const int tstride = N / (NS * R);
  cx<T> v[QM][R];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int idx = tid + q * NT;
    if (idx < total) {
      const int j = idx / count, t = idx - j * count;
      const int b = base(t);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = j + r * NB;
        v[q][r] = (first && !hk.live_in(t, e)) ? mk<T>(0, 0) : buf[b + e * es];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int idx = tid + q * NT;
    if (idx < total) {
      const int j = idx / count, t = idx - j * count;
      const int b = base(t);
      const int k = j % NS;
      if (NS > 1) {
        const int ks = k * tstride;
        int ti = 0;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          ti += ks;
          cx<T> w = tw[ti];
          if constexpr (SIGN < 0) w.y = -w.y;
          v[q][r] = v[q][r] * w;
        }
      }
      dft<T, R, SIGN>(v[q]);
      const int o = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = o + r * NS;
        cx<T> x = v[q][r];
        if (last) {
          if (!hk.live_out(t, e)) continue;
          hk.post(t, e, x);
        }
        buf[b + e * es] = x;
      }
    }
  }
  __syncthreads();
}

// A stage of prime radix P without a register butterfly: output ro of butterfly j is
// sum_r x[j + r N/P] w_N^(r m), m = (N / (NS P)) (k + ro NS), k = j mod NS (the stage
// twiddle and the length-P DFT in one exponent).
template <typename T, int SIGN, int NT, typename BaseF, typename Hooks>
__device__ __forceinline__ void rt_stage_prime(cx<T>* buf, int N, int NS, int P, bool first, bool last, int count,
                                               int es, BaseF base, const cx<T>* tw, int tid, const Hooks& hk) {
  const int NB = N / P;
  const int total = count * N;
  const int tstride = N / (NS * P);
  cx<T> acc[kRtEmax];
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) {
    const int idx = tid + q * NT;
    acc[q] = mk<T>(0, 0);
    if (idx < total) {
      const int rest = idx / count, t = idx - rest * count;
      const int ro = rest / NB, j = rest - ro * NB;
      const int b = base(t);
      const int m = tstride * (j % NS + ro * NS);
      int ti = 0;
      cx<T> s = mk<T>(0, 0);
      for (int r = 0; r < P; ++r) {
        const int e = j + r * NB;
        if (!first || hk.live_in(t, e)) {
          cx<T> w = tw[ti];
          if constexpr (SIGN < 0) w.y = -w.y;
          s = s + buf[b + e * es] * w;
        }
        ti += m;
        if (ti >= N) ti -= N;
      }
      acc[q] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) {
    const int idx = tid + q * NT;
    if (idx < total) {
      const int rest = idx / count, t = idx - rest * count;
      const int ro = rest / NB, j = rest - ro * NB;
      const int b = base(t);
      const int k = j % NS;
      const int e = (j / NS) * NS * P + k + ro * NS;
      cx<T> x = acc[q];
      if (last) {
        if (!hk.live_out(t, e)) continue;
        hk.post(t, e, x);
      }
      buf[b + e * es] = x;
    }
  }
  __syncthreads();
}

template <typename T, int SIGN, int NT, typename BaseF, typename Hooks = NoHooks>
__device__ __forceinline__ void rt_fft(cx<T>* buf, const FftPlan& p, int count, int es, BaseF base, const cx<T>* tw, int tid,
                       const Hooks& hk = Hooks()) {
  int ns = 1;
  for (int st = 0; st < p.nst; ++st) {
    const int R = p.radix[st];
    const bool first = st == 0, last = st == p.nst - 1;
    switch (R) {
      case 2: rt_stage<T, 2, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      case 3: rt_stage<T, 3, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      case 4: rt_stage<T, 4, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      case 5: rt_stage<T, 5, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      case 7: rt_stage<T, 7, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      case 8: rt_stage<T, 8, SIGN, NT>(buf, p.n, ns, first, last, count, es, base, tw, tid, hk); break;
      default: rt_stage_prime<T, SIGN, NT>(buf, p.n, ns, R, first, last, count, es, base, tw, tid, hk); break;
    }
    ns *= R;
  }
}

// odd row stride: a wave of consecutive transforms touches distinct banks
__host__ __device__ inline int odd_stride(int n) { return (n & 1) ? n : n + 1; }

// ---------------------------------------------------------------------------
// i3 passes over the sphere columns (same data movement as k_zinv / k_zfwd)
// ---------------------------------------------------------------------------
constexpr int kRtColThreads = 256;

// Block (cblk, b) of the i3 passes: columns cblk*CT .. +CT-1 of band b.  The bodies take
// the block coordinates as arguments so the grouped kernels below run the same code.
template <typename T, int NT>
__device__ __forceinline__ void zinv_rt_block(const SphereMaps& m, int CT, const cx<T>* __restrict__ tw3g,
                                              const cx<T>* __restrict__ X, int ldx, cx<T>* __restrict__ W,
                                              int cblk, int b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NZ = m.nr3;
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw = tile + NZ * CT;
  const int tid = threadIdx.x;
  const int c0 = cblk * CT;
  const int cn = min(CT, m.ncol_pad - c0);
  const int kb = m.colstart[c0], ke = m.colstart[c0 + cn];
  const cx<T>* xb = X + (size_t)b * ldx;
  for (int i = tid; i < NZ; i += NT) tw[i] = tw3g[i];
  for (int i = tid; i < NZ * CT; i += NT) tile[i] = mk<T>(0, 0);
  __syncthreads();
  for (int k = kb + tid; k < ke; k += NT) tile[m.kz[k] * CT + (m.kcol[k] - c0)] = xb[k];
  __syncthreads();
  rt_fft<T, +1, NT>(tile, m.plan[2], CT, CT, [](int t) { return t; }, tw, tid);
  cx<T>* wb = W + (size_t)b * NZ * m.ncol_pad + c0;
  for (int i = tid; i < NZ * CT; i += NT) {
    const int z = i / CT, c = i - z * CT;
    if (c < cn) wb[(size_t)z * m.ncol_pad + c] = tile[i];
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_zinv_rt(SphereMaps m, int CT, const cx<T>* __restrict__ tw3g,
                                                const cx<T>* __restrict__ X, int ldx, cx<T>* __restrict__ W) {
  zinv_rt_block<T, NT>(m, CT, tw3g, X, ldx, W, blockIdx.x, blockIdx.y);
}

// out = alpha (FFT/nnr + kin X + Hnl) + beta X + gamma Z on the sphere entries
template <typename T, int NT>
__device__ __forceinline__ void zfwd_rt_block(const SphereMaps& m, int CT, const cx<T>* __restrict__ tw3g,
                                              const cx<T>* __restrict__ W, const T* __restrict__ kin,
                                              const cx<T>* __restrict__ X, const cx<T>* __restrict__ Hnl,
                                              const cx<T>* __restrict__ Z, cx<T>* __restrict__ out, int ld,
                                              T inv_nnr, T alpha, T beta, T gamma, int cblk, int b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NZ = m.nr3;
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw = tile + NZ * CT;
  const int tid = threadIdx.x;
  const int c0 = cblk * CT;
  const int cn = min(CT, m.ncol_pad - c0);
  const int kb = m.colstart[c0], ke = m.colstart[c0 + cn];
  const cx<T>* wb = W + (size_t)b * NZ * m.ncol_pad + c0;
  for (int i = tid; i < NZ; i += NT) tw[i] = tw3g[i];
  for (int i = tid; i < NZ * CT; i += NT) {
    const int z = i / CT, c = i - z * CT;
    tile[i] = c < cn ? wb[(size_t)z * m.ncol_pad + c] : mk<T>(0, 0);
  }
  __syncthreads();
  rt_fft<T, -1, NT>(tile, m.plan[2], CT, CT, [](int t) { return t; }, tw, tid);
  const size_t ob = (size_t)b * ld;
  for (int k = kb + tid; k < ke; k += NT) {
    const cx<T> x = X[ob + k];
    cx<T> h = inv_nnr * tile[m.kz[k] * CT + (m.kcol[k] - c0)];
    h = h + kin[k] * x;
    if (Hnl) h = h + Hnl[ob + k];
    cx<T> o = alpha * h;
    if (beta != T(0)) o = o + beta * x;
    if (Z) o = o + gamma * Z[ob + k];
    out[ob + k] = o;
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_zfwd_rt(SphereMaps m, int CT, const cx<T>* __restrict__ tw3g,
                                                const cx<T>* __restrict__ W, const T* __restrict__ kin,
                                                const cx<T>* __restrict__ X, const cx<T>* __restrict__ Hnl,
                                                const cx<T>* __restrict__ Z, cx<T>* __restrict__ out, int ld,
                                                T inv_nnr, T alpha, T beta, T gamma) {
  zfwd_rt_block<T, NT>(m, CT, tw3g, W, kin, X, Hnl, Z, out, ld, inv_nnr, alpha, beta, gamma, blockIdx.x,
                       blockIdx.y);
}

// ---------------------------------------------------------------------------
// Plane pass in LDS (any n1 x n2): inverse i2 on the active i1 columns, inverse i1 on
// every row times V, forward i1 keeping the active columns, forward i2
// ---------------------------------------------------------------------------
template <typename T>
static size_t plane_rt_smem(int n1, int n2) {
  return sizeof(cx<T>) * ((size_t)n2 * odd_stride(n1) + n1 + n2) + sizeof(int) * n1 + n1;
}

template <typename T, int NT>
__device__ __forceinline__ void plane_rt_block(const SphereMaps& m, const cx<T>* __restrict__ tw1g,
                                               const cx<T>* __restrict__ tw2g, const T* __restrict__ V,
                                               cx<T>* __restrict__ W, int z, int b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N1 = m.nr1, N2 = m.nr2, S1 = odd_stride(N1);
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw1 = tile + (size_t)N2 * S1;
  cx<T>* tw2 = tw1 + N1;
  int* xa = reinterpret_cast<int*>(tw2 + N2);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int nxa = m.nxa;
  for (int i = tid; i < N1; i += NT) { tw1[i] = tw1g[i]; xact[i] = 0; }
  for (int i = tid; i < N2; i += NT) tw2[i] = tw2g[i];
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  for (int i = tid; i < nxa * N2; i += NT) {
    const int y = i / nxa, c = i - y * nxa;
    tile[y * S1 + m.xa[c]] = mk<T>(0, 0);
  }
  __syncthreads();
  cx<T>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  for (int c = tid; c < m.ncol; c += NT) {
    const int xy = m.colxy[c];
    tile[(xy / N1) * S1 + xy % N1] = wrow[c];
  }
  __syncthreads();
  auto col_base = [&](int t) { return xa[t]; };
  auto row_base = [&](int t) { return t * S1; };
  rt_fft<T, +1, NT>(tile, m.plan[1], nxa, S1, col_base, tw2, tid);
  const XInvHooks<T> hinv{xact, V + (size_t)z * N1 * N2, N1};
  rt_fft<T, +1, NT>(tile, m.plan[0], N2, 1, row_base, tw1, tid, hinv);
  const XFwdHooks hfwd{xact};
  rt_fft<T, -1, NT>(tile, m.plan[0], N2, 1, row_base, tw1, tid, hfwd);
  rt_fft<T, -1, NT>(tile, m.plan[1], nxa, S1, col_base, tw2, tid);
  for (int c = tid; c < m.ncol; c += NT) {
    const int xy = m.colxy[c];
    wrow[c] = tile[(xy / N1) * S1 + xy % N1];
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_plane_rt(SphereMaps m, const cx<T>* __restrict__ tw1g,
                                                 const cx<T>* __restrict__ tw2g, const T* __restrict__ V,
                                                 cx<T>* __restrict__ W) {
  plane_rt_block<T, NT>(m, tw1g, tw2g, V, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Grouped passes: one launch covers the H|psi> pass of every system of a batch that runs
// on runtime plans.  Block L of the 1D grid belongs to the system whose first block blk0 is
// the largest one <= L; the system's descriptor carries its maps, plans, constants,
// buffers and epilogue scalars, and the block runs the same body as the one-system kernels
// above.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int group_find(const GroupHpsi<T>* __restrict__ g, int n, int L) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].blk0 <= L) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_zinv_group(const GroupHpsi<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupHpsi<T>& p = g[group_find(g, n, L)];
  const int local = L - p.blk0, ncb = p.m.ncol_pad / p.ct;
  zinv_rt_block<T, NT>(p.m, p.ct, p.tw3, p.X, p.ld, p.W, local % ncb, local / ncb);
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_zfwd_group(const GroupHpsi<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupHpsi<T>& p = g[group_find(g, n, L)];
  const int local = L - p.blk0, ncb = p.m.ncol_pad / p.ct;
  zfwd_rt_block<T, NT>(p.m, p.ct, p.tw3, p.W, p.kin, p.X, p.Hnl, p.Z, p.out, p.ld, p.inv_nnr, p.alpha, p.beta,
                       p.gamma, local % ncb, local / ncb);
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_plane_group(const GroupHpsi<T>* __restrict__ g, int n) {
  const int L = blockIdx.x;
  const GroupHpsi<T>& p = g[group_find(g, n, L)];
  const int local = L - p.blk0;
  plane_rt_block<T, NT>(p.m, p.tw1, p.tw2, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

// ---------------------------------------------------------------------------
// Split plane pass: Y[bl][i3][a][i2] holds the i2 lines of the active i1 columns a
// ---------------------------------------------------------------------------
// SIGN = +1: scatter the sphere columns of active columns g0 .. g0+XG-1 into lines,
// inverse i2, write Y.  SIGN = -1: read Y, forward i2, write the sphere columns back to W.
template <typename T, int SIGN, int NT>
__global__ __launch_bounds__(NT) void k_ysplit_rt(SphereMaps m, int XG, const cx<T>* __restrict__ tw2g,
                                                  cx<T>* __restrict__ W, cx<T>* __restrict__ Y, int band0) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N1 = m.nr1, N2 = m.nr2, S2 = odd_stride(N2), nxa = m.nxa;
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw = tile + (size_t)XG * S2;
  const int tid = threadIdx.x;
  const int g0 = blockIdx.x * XG, z = blockIdx.y, bl = blockIdx.z;
  const int nl = min(XG, nxa - g0);
  cx<T>* wrow = W + ((size_t)(band0 + bl) * m.nr3 + z) * m.ncol_pad;
  cx<T>* yl = Y + (((size_t)bl * m.nr3 + z) * nxa + g0) * N2;
  const int cs = m.xcstart[g0], ce = m.xcstart[g0 + nl];
  for (int i = tid; i < N2; i += NT) tw[i] = tw2g[i];
  if constexpr (SIGN > 0) {
    for (int i = tid; i < XG * S2; i += NT) tile[i] = mk<T>(0, 0);
    __syncthreads();
    for (int k = cs + tid; k < ce; k += NT) {
      const int c = m.xcols[k];
      tile[(m.colxa[c] - g0) * S2 + m.colxy[c] / N1] = wrow[c];
    }
  } else {
    for (int i = tid; i < XG * N2; i += NT) {
      const int li = i / N2, y = i - li * N2;
      tile[li * S2 + y] = li < nl ? yl[(size_t)li * N2 + y] : mk<T>(0, 0);
    }
  }
  __syncthreads();
  rt_fft<T, SIGN, NT>(tile, m.plan[1], XG, 1, [&](int t) { return t * S2; }, tw, tid);
  if constexpr (SIGN > 0) {
    for (int i = tid; i < nl * N2; i += NT) {
      const int li = i / N2, y = i - li * N2;
      yl[(size_t)li * N2 + y] = tile[li * S2 + y];
    }
  } else {
    for (int k = cs + tid; k < ce; k += NT) {
      const int c = m.xcols[k];
      wrow[c] = tile[(m.colxa[c] - g0) * S2 + m.colxy[c] / N1];
    }
  }
}

// V multiplication of the split row pass; rows at or beyond nrows (the partial last group)
// are padding and become zero
template <typename T>
struct XSplitInvHooks {
  const unsigned char* act;
  const T* vrow;   // V of the group's first row
  int n1, nrows;
  __device__ __forceinline__ bool live_in(int, int e) const { return act[e] != 0; }
  __device__ __forceinline__ bool live_out(int, int) const { return true; }
  __device__ __forceinline__ void post(int t, int e, cx<T>& v) const {
    v = t < nrows ? vrow[t * n1 + e] * v : mk<T>(0, 0);
  }
};

// rows i2 = y0 .. y0+YG-1 of plane i3: inverse i1, times V, forward i1, active columns back
// to Y
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_xsplit_rt(SphereMaps m, int YG, const cx<T>* __restrict__ tw1g,
                                                  const T* __restrict__ V, cx<T>* __restrict__ Y) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N1 = m.nr1, N2 = m.nr2, S1 = odd_stride(N1), nxa = m.nxa;
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw = tile + (size_t)YG * S1;
  int* xa = reinterpret_cast<int*>(tw + N1);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * YG, z = blockIdx.y, bl = blockIdx.z;
  const int ny = min(YG, N2 - y0);
  cx<T>* yb = Y + ((size_t)bl * m.nr3 + z) * nxa * N2;
  for (int i = tid; i < N1; i += NT) { tw[i] = tw1g[i]; xact[i] = 0; }
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  __syncthreads();
  for (int i = tid; i < nxa * YG; i += NT) {
    const int a = i / YG, li = i - a * YG;
    tile[li * S1 + xa[a]] = li < ny ? yb[(size_t)a * N2 + y0 + li] : mk<T>(0, 0);
  }
  __syncthreads();
  auto row_base = [&](int t) { return t * S1; };
  const XSplitInvHooks<T> hinv{xact, V + ((size_t)z * N2 + y0) * N1, N1, ny};
  rt_fft<T, +1, NT>(tile, m.plan[0], YG, 1, row_base, tw, tid, hinv);
  const XFwdHooks hfwd{xact};
  rt_fft<T, -1, NT>(tile, m.plan[0], YG, 1, row_base, tw, tid, hfwd);
  for (int i = tid; i < nxa * YG; i += NT) {
    const int a = i / YG, li = i - a * YG;
    if (li < ny) yb[(size_t)a * N2 + y0 + li] = tile[li * S1 + xa[a]];
  }
}

// ---------------------------------------------------------------------------
// Density (EvalRhoPsi.cpp:138-150) with runtime plans
// ---------------------------------------------------------------------------
// part[g][i3][i2][i1] = sum over bands b = g, g + groups, ... of wts[b] |psi_b|^2
template <int NT>
__global__ __launch_bounds__(NT) void k_plane_rho_rt(SphereMaps m, const cx<double>* __restrict__ tw1g,
                                                     const cx<double>* __restrict__ tw2g,
                                                     const cx<double>* __restrict__ W, int nb, int groups,
                                                     const double* __restrict__ wts, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N1 = m.nr1, N2 = m.nr2, S1 = odd_stride(N1), nxa = m.nxa;
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  cx<double>* tw1 = tile + (size_t)N2 * S1;
  cx<double>* tw2 = tw1 + N1;
  int* xa = reinterpret_cast<int*>(tw2 + N2);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int z = blockIdx.x, g0 = blockIdx.y;
  const int npts = N1 * N2;
  for (int i = tid; i < N1; i += NT) { tw1[i] = tw1g[i]; xact[i] = 0; }
  for (int i = tid; i < N2; i += NT) tw2[i] = tw2g[i];
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  double acc[kRtEmax];
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) acc[q] = 0.0;
  const XLiveHooks hk{xact};
  auto col_base = [&](int t) { return xa[t]; };
  auto row_base = [&](int t) { return t * S1; };
  for (int b = g0; b < nb; b += groups) {
    __syncthreads();
    for (int i = tid; i < nxa * N2; i += NT) {
      const int y = i / nxa, c = i - y * nxa;
      tile[y * S1 + m.xa[c]] = mk<double>(0, 0);
    }
    __syncthreads();
    const cx<double>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
    for (int c = tid; c < m.ncol; c += NT) {
      const int xy = m.colxy[c];
      tile[(xy / N1) * S1 + xy % N1] = wrow[c];
    }
    __syncthreads();
    rt_fft<double, +1, NT>(tile, m.plan[1], nxa, S1, col_base, tw2, tid);
    rt_fft<double, +1, NT>(tile, m.plan[0], N2, 1, row_base, tw1, tid, hk);
    const double wb = wts[b];
#pragma unroll
    for (int q = 0; q < kRtEmax; ++q) {
      const int idx = tid + q * NT;
      if (idx < npts) {
        const int y = idx / N1, x = idx - y * N1;
        const cx<double> v = tile[y * S1 + x];
        acc[q] += wb * (v.x * v.x + v.y * v.y);
      }
    }
  }
  double* out = part + ((size_t)g0 * m.nr3 + z) * (size_t)npts;
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) {
    const int idx = tid + q * NT;
    if (idx < npts) out[idx] = acc[q];
  }
}

// rows y0 .. y0+YG-1 of plane i3 for the nbc bands of a chunk: inverse i1 of Y's active
// columns, rho (+)= sum_b wts[b] |psi_b|^2 (accumulate: add to what an earlier chunk wrote)
template <int NT>
__global__ __launch_bounds__(NT) void k_xrho_split(SphereMaps m, int YG, const cx<double>* __restrict__ tw1g,
                                                   const cx<double>* __restrict__ Y, int nbc,
                                                   const double* __restrict__ wts, double* __restrict__ rho,
                                                   int accumulate) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N1 = m.nr1, N2 = m.nr2, S1 = odd_stride(N1), nxa = m.nxa;
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  cx<double>* tw = tile + (size_t)YG * S1;
  int* xa = reinterpret_cast<int*>(tw + N1);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * YG, z = blockIdx.y;
  const int ny = min(YG, N2 - y0);
  const int npts = YG * N1;
  for (int i = tid; i < N1; i += NT) { tw[i] = tw1g[i]; xact[i] = 0; }
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  double acc[kRtEmax];
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) acc[q] = 0.0;
  const XLiveHooks hk{xact};
  auto row_base = [&](int t) { return t * S1; };
  for (int bl = 0; bl < nbc; ++bl) {
    __syncthreads();
    const cx<double>* yb = Y + ((size_t)bl * m.nr3 + z) * nxa * N2;
    for (int i = tid; i < nxa * YG; i += NT) {
      const int a = i / YG, li = i - a * YG;
      tile[li * S1 + xa[a]] = li < ny ? yb[(size_t)a * N2 + y0 + li] : mk<double>(0, 0);
    }
    __syncthreads();
    rt_fft<double, +1, NT>(tile, m.plan[0], YG, 1, row_base, tw, tid, hk);
    const double wb = wts[bl];
#pragma unroll
    for (int q = 0; q < kRtEmax; ++q) {
      const int idx = tid + q * NT;
      if (idx < npts) {
        const int li = idx / N1, x = idx - li * N1;
        const cx<double> v = tile[li * S1 + x];
        acc[q] += wb * (v.x * v.x + v.y * v.y);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kRtEmax; ++q) {
    const int idx = tid + q * NT;
    if (idx < npts) {
      const int li = idx / N1, x = idx - li * N1;
      if (li < ny) {
        double* r = rho + ((size_t)z * N2 + y0 + li) * N1 + x;
        *r = (accumulate ? *r : 0.0) + acc[q];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dense lines (fft_fftw3.cpp:13-46 one axis at a time): array [outer][N][inner], CT lines
// per block as k_fft_lines
// ---------------------------------------------------------------------------
template <int SIGN, int NT>
__global__ __launch_bounds__(NT) void k_lines_rt(cx<double>* __restrict__ data, FftPlan p, int inner, int nlines,
                                                 int CT, const cx<double>* __restrict__ twg, double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = p.n, S = CT + 1;
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  cx<double>* tw = tile + (size_t)N * S;
  const int tid = threadIdx.x;
  const long long l0 = (long long)blockIdx.x * CT;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  auto addr = [&](int t, int e) -> long long {
    const long long L = l0 + t;
    if (inner == 1) return L * N + e;
    const long long o = L / inner, i = L - o * inner;
    return o * N * inner + i + (long long)e * inner;
  };
  for (int idx = tid; idx < N * CT; idx += NT) {
    int t, e;
    if (inner == 1) { t = idx / N; e = idx - t * N; }
    else { e = idx / CT; t = idx - e * CT; }
    tile[e * S + t] = (l0 + t < nlines) ? data[addr(t, e)] : mk<double>(0, 0);
  }
  __syncthreads();
  rt_fft<double, SIGN, NT>(tile, p, CT, S, [](int t) { return t; }, tw, tid);
  for (int idx = tid; idx < N * CT; idx += NT) {
    int t, e;
    if (inner == 1) { t = idx / N; e = idx - t * N; }
    else { e = idx / CT; t = idx - e * CT; }
    if (l0 + t < nlines) data[addr(t, e)] = scale * tile[e * S + t];
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
constexpr size_t kLdsMax = 160 * 1024;

template <typename K>
static hipError_t allow_lds(K kernel, size_t bytes) {
  if (bytes > kLdsMax) return hipErrorInvalidValue;
  if (bytes <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kLdsMax);
}

// lines per block: the largest power of two <= 16 whose transforms fit the block
static int rt_lines_per_block(const FftPlan& p, int nt, int most = 16) {
  for (int ct = most; ct >= 1; ct /= 2)
    if (rt_fits(p, ct, nt)) return ct;
  return 0;
}

template <typename T>
hipError_t launch_zinv_rt(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W, int nb,
                          hipStream_t s) {
  constexpr int NT = kRtColThreads;
  const int ct = rt_lines_per_block(m.plan[2], NT);
  if (!ct || m.plan[2].n != m.nr3) return hipErrorInvalidValue;
  const size_t smem = sizeof(cx<T>) * ((size_t)m.nr3 * ct + m.nr3);
  hipError_t e = allow_lds(&k_zinv_rt<T, NT>, smem);
  if (e != hipSuccess) return e;
  dim3 grid(m.ncol_pad / ct, nb);
  hipLaunchKernelGGL((k_zinv_rt<T, NT>), grid, dim3(NT), smem, s, m, ct, static_cast<const cx<T>*>(tw3), X, ldx,
                     W);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zfwd_rt(const SphereMaps& m, const void* tw3, const cx<T>* W, const T* kin, const cx<T>* X,
                          const cx<T>* Hnl, const cx<T>* Z, cx<T>* out, int ld, int nb, T inv_nnr, T alpha,
                          T beta, T gamma, hipStream_t s) {
  constexpr int NT = kRtColThreads;
  const int ct = rt_lines_per_block(m.plan[2], NT);
  if (!ct || m.plan[2].n != m.nr3) return hipErrorInvalidValue;
  const size_t smem = sizeof(cx<T>) * ((size_t)m.nr3 * ct + m.nr3);
  hipError_t e = allow_lds(&k_zfwd_rt<T, NT>, smem);
  if (e != hipSuccess) return e;
  dim3 grid(m.ncol_pad / ct, nb);
  hipLaunchKernelGGL((k_zfwd_rt<T, NT>), grid, dim3(NT), smem, s, m, ct, static_cast<const cx<T>*>(tw3), W, kin,
                     X, Hnl, Z, out, ld, inv_nnr, alpha, beta, gamma);
  return hipGetLastError();
}

// lanes of the LDS plane kernel: 256 when the plane's transforms fit, else 1024
static int plane_rt_threads(const SphereMaps& m) {
  for (int nt : {256, 1024})
    if (rt_fits(m.plan[0], m.nr2, nt) && rt_fits(m.plan[1], m.nr1, nt)) return nt;
  return 0;
}

bool plane_rt_fits(int n1, int n2, bool fp64) {
  FftPlan p1, p2;
  if (!fft_plan(n1, &p1) || !fft_plan(n2, &p2)) return false;
  if (!rt_fits(p1, n2, 1024) || !rt_fits(p2, n1, 1024)) return false;
  return (fp64 ? plane_rt_smem<double>(n1, n2) : plane_rt_smem<float>(n1, n2)) <= kLdsMax;
}

size_t plane_split_elems(const SphereMaps& m, int nbc) {
  return (size_t)nbc * m.nr3 * m.nxa * m.nr2;
}

int plane_split_chunk(const SphereMaps& m, int nb, size_t bytes_per_complex, size_t budget_bytes) {
  const size_t per_band = plane_split_elems(m, 1) * bytes_per_complex;
  const size_t c = per_band ? budget_bytes / per_band : (size_t)nb;
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)nb, c));
}

template <typename T>
static bool use_split(const SphereMaps& m) {
  return m.path == FFT_SPLIT || !plane_rt_fits(m.nr1, m.nr2, sizeof(T) == sizeof(double));
}

template <typename T>
static hipError_t split_shapes(const SphereMaps& m, int* xg, int* yg, size_t* ysm, size_t* xsm) {
  *xg = rt_lines_per_block(m.plan[1], kRtColThreads);
  *yg = rt_lines_per_block(m.plan[0], kRtColThreads);
  if (!*xg || !*yg) return hipErrorInvalidValue;
  *ysm = sizeof(cx<T>) * ((size_t)*xg * odd_stride(m.nr2) + m.nr2);
  *xsm = sizeof(cx<T>) * ((size_t)*yg * odd_stride(m.nr1) + m.nr1) + sizeof(int) * m.nr1 + m.nr1;
  return hipSuccess;
}

template <typename T>
hipError_t launch_plane_rt(const SphereMaps& m, const void* tw1, const void* tw2, const T* V, cx<T>* W, int nb,
                           cx<T>* Y, int chunk, hipStream_t s) {
  if (m.plan[0].n != m.nr1 || m.plan[1].n != m.nr2) return hipErrorInvalidValue;
  const cx<T>* t1 = static_cast<const cx<T>*>(tw1);
  const cx<T>* t2 = static_cast<const cx<T>*>(tw2);
  if (!use_split<T>(m)) {
    const int nt = plane_rt_threads(m);
    const size_t smem = plane_rt_smem<T>(m.nr1, m.nr2);
    dim3 grid(m.nr3, nb);
    hipError_t e;
    if (nt == 256) {
      if ((e = allow_lds(&k_plane_rt<T, 256>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_rt<T, 256>), grid, dim3(256), smem, s, m, t1, t2, V, W);
    } else if (nt == 1024) {
      if ((e = allow_lds(&k_plane_rt<T, 1024>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_rt<T, 1024>), grid, dim3(1024), smem, s, m, t1, t2, V, W);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (!Y || chunk < 1) return hipErrorInvalidValue;
  int xg, yg;
  size_t ysm, xsm;
  hipError_t e = split_shapes<T>(m, &xg, &yg, &ysm, &xsm);
  if (e != hipSuccess) return e;
  if ((e = allow_lds(&k_ysplit_rt<T, +1, kRtColThreads>, ysm)) != hipSuccess ||
      (e = allow_lds(&k_ysplit_rt<T, -1, kRtColThreads>, ysm)) != hipSuccess ||
      (e = allow_lds(&k_xsplit_rt<T, kRtColThreads>, xsm)) != hipSuccess)
    return e;
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = std::min(chunk, nb - b0);
    const dim3 gy((m.nxa + xg - 1) / xg, m.nr3, nbc), gx((m.nr2 + yg - 1) / yg, m.nr3, nbc);
    hipLaunchKernelGGL((k_ysplit_rt<T, +1, kRtColThreads>), gy, dim3(kRtColThreads), ysm, s, m, xg, t2, W, Y, b0);
    hipLaunchKernelGGL((k_xsplit_rt<T, kRtColThreads>), gx, dim3(kRtColThreads), xsm, s, m, yg, t1, V, Y);
    hipLaunchKernelGGL((k_ysplit_rt<T, -1, kRtColThreads>), gy, dim3(kRtColThreads), ysm, s, m, xg, t2, W, Y, b0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
bool group_fft_shape(const SphereMaps& m, int nb, int* ct, int* zblocks, int* pblocks, size_t* zsmem,
                     size_t* psmem) {
  constexpr int NT = kRtColThreads;
  if (m.plan[0].n != m.nr1 || m.plan[1].n != m.nr2 || m.plan[2].n != m.nr3) return false;
  if (use_split<T>(m) || plane_rt_threads(m) != NT) return false;
  *ct = rt_lines_per_block(m.plan[2], NT);
  if (!*ct) return false;
  *zblocks = m.ncol_pad / *ct * nb;
  *pblocks = m.nr3 * nb;
  *zsmem = sizeof(cx<T>) * ((size_t)m.nr3 * *ct + m.nr3);
  *psmem = plane_rt_smem<T>(m.nr1, m.nr2);
  return *zsmem <= kLdsMax && *psmem <= kLdsMax;
}

template <typename T>
hipError_t launch_group_fft_rt(int pass, const GroupHpsi<T>* g, int n, int blocks, size_t smem, hipStream_t s) {
  constexpr int NT = kRtColThreads;
  if (n < 1 || blocks < 1) return hipErrorInvalidValue;
  hipError_t e;
  switch (pass) {
    case 0:
      if ((e = allow_lds(&k_zinv_group<T, NT>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_zinv_group<T, NT>), dim3(blocks), dim3(NT), smem, s, g, n);
      break;
    case 1:
      if ((e = allow_lds(&k_plane_group<T, NT>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_group<T, NT>), dim3(blocks), dim3(NT), smem, s, g, n);
      break;
    case 2:
      if ((e = allow_lds(&k_zfwd_group<T, NT>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_zfwd_group<T, NT>), dim3(blocks), dim3(NT), smem, s, g, n);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ void k_sum_groups_rt(const double* __restrict__ part, int groups, size_t n, double* __restrict__ rho) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int g = 0; g < groups; ++g) s += part[(size_t)g * n + i];
  rho[i] = s;
}

hipError_t launch_plane_rho_rt(const SphereMaps& m, const void* tw1, const void* tw2, const cx<double>* W, int nb,
                               int groups, const double* wts, double* part, double* rho, cx<double>* Y, int chunk,
                               hipStream_t s) {
  if (m.plan[0].n != m.nr1 || m.plan[1].n != m.nr2 || groups < 1) return hipErrorInvalidValue;
  const cx<double>* t1 = static_cast<const cx<double>*>(tw1);
  const cx<double>* t2 = static_cast<const cx<double>*>(tw2);
  const int npts = m.nr1 * m.nr2;
  if (!use_split<double>(m)) {
    // 256 lanes when the plane's points and transforms fit them, else 1024
    int nt = 0;
    for (int c : {256, 1024})
      if (!nt && npts <= c * kRtEmax && rt_fits(m.plan[0], m.nr2, c) && rt_fits(m.plan[1], m.nr1, c)) nt = c;
    const size_t smem = plane_rt_smem<double>(m.nr1, m.nr2);
    hipError_t e;
    if (nt == 256) {
      if ((e = allow_lds(&k_plane_rho_rt<256>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_rho_rt<256>), dim3(m.nr3, groups), dim3(256), smem, s, m, t1, t2, W, nb, groups,
                         wts, part);
    } else if (nt == 1024) {
      if ((e = allow_lds(&k_plane_rho_rt<1024>, smem)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_rho_rt<1024>), dim3(m.nr3, groups), dim3(1024), smem, s, m, t1, t2, W, nb,
                         groups, wts, part);
    }
    if (nt) {
      const size_t n = (size_t)m.nnr;
      hipLaunchKernelGGL(k_sum_groups_rt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, groups, n, rho);
      return hipGetLastError();
    }
  }
  // split: i2 lines into Y per chunk of bands, then the rows accumulate into rho
  if (!Y || chunk < 1) return hipErrorInvalidValue;
  int xg, yg;
  size_t ysm, xsm;
  hipError_t e = split_shapes<double>(m, &xg, &yg, &ysm, &xsm);
  if (e != hipSuccess) return e;
  if ((e = allow_lds(&k_ysplit_rt<double, +1, kRtColThreads>, ysm)) != hipSuccess ||
      (e = allow_lds(&k_xrho_split<kRtColThreads>, xsm)) != hipSuccess)
    return e;
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = std::min(chunk, nb - b0);
    const dim3 gy((m.nxa + xg - 1) / xg, m.nr3, nbc), gx((m.nr2 + yg - 1) / yg, m.nr3);
    hipLaunchKernelGGL((k_ysplit_rt<double, +1, kRtColThreads>), gy, dim3(kRtColThreads), ysm, s, m, xg, t2,
                       const_cast<cx<double>*>(W), Y, b0);
    hipLaunchKernelGGL((k_xrho_split<kRtColThreads>), gx, dim3(kRtColThreads), xsm, s, m, yg, t1, Y, nbc, wts + b0,
                       rho, b0 > 0 ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_fft_lines_rt(cx<double>* data, const FftPlan& p, int inner, int nlines, const cx<double>* tw,
                               int sign, double scale, hipStream_t s) {
  constexpr int NT = kRtColThreads;
  const int ct = rt_lines_per_block(p, NT);
  if (!ct || nlines < 1 || inner < 1) return hipErrorInvalidValue;
  const size_t smem = sizeof(cx<double>) * ((size_t)p.n * (ct + 1) + p.n);
  hipError_t e = allow_lds(sign > 0 ? &k_lines_rt<+1, NT> : &k_lines_rt<-1, NT>, smem);
  if (e != hipSuccess) return e;
  dim3 grid((nlines + ct - 1) / ct);
  if (sign > 0)
    hipLaunchKernelGGL((k_lines_rt<+1, NT>), grid, dim3(NT), smem, s, data, p, inner, nlines, ct, tw, scale);
  else
    hipLaunchKernelGGL((k_lines_rt<-1, NT>), grid, dim3(NT), smem, s, data, p, inner, nlines, ct, tw, scale);
  return hipGetLastError();
}

template bool group_fft_shape<double>(const SphereMaps&, int, int*, int*, int*, size_t*, size_t*);
template bool group_fft_shape<float>(const SphereMaps&, int, int*, int*, int*, size_t*, size_t*);
template hipError_t launch_group_fft_rt<double>(int, const GroupHpsi<double>*, int, int, size_t, hipStream_t);
template hipError_t launch_group_fft_rt<float>(int, const GroupHpsi<float>*, int, int, size_t, hipStream_t);
template hipError_t launch_zinv_rt<double>(const SphereMaps&, const void*, const cx<double>*, int, cx<double>*, int,
                                           hipStream_t);
template hipError_t launch_zinv_rt<float>(const SphereMaps&, const void*, const cx<float>*, int, cx<float>*, int,
                                          hipStream_t);
template hipError_t launch_zfwd_rt<double>(const SphereMaps&, const void*, const cx<double>*, const double*,
                                           const cx<double>*, const cx<double>*, const cx<double>*, cx<double>*, int,
                                           int, double, double, double, double, hipStream_t);
template hipError_t launch_zfwd_rt<float>(const SphereMaps&, const void*, const cx<float>*, const float*,
                                          const cx<float>*, const cx<float>*, const cx<float>*, cx<float>*, int, int,
                                          float, float, float, float, hipStream_t);
template hipError_t launch_plane_rt<double>(const SphereMaps&, const void*, const void*, const double*, cx<double>*,
                                            int, cx<double>*, int, hipStream_t);
template hipError_t launch_plane_rt<float>(const SphereMaps&, const void*, const void*, const float*, cx<float>*, int,
                                           cx<float>*, int, hipStream_t);

}  // namespace pwh
