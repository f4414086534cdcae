// pwh_fft.hip -- the local part of H|psi> on gfx950: a sphere-pruned 3D FFT pair with
// the local potential fused into the real-space pass and the kinetic term (plus the
// nonlocal result and the Chebyshev three-term update) fused into the last pass.
//
// Replaces, for a whole block of bands at once, the per-band sequence of
// ApplyHam_block.cpp:33-61: scatter through N123, fft_fftw3(inverse), *rho,
// fft_fftw3(forward, 1/N), gather + 0.5*XKG*TPIBA2.
//
// Three passes per block of nb bands:
//   zinv   (grid ncol_pad/CT x nb): gather each column's sphere entries, inverse
//          FFT along i3 in LDS, write W[b][i3][col].
//   plane  (grid nr3 x nb): one (i3, band) plane per workgroup, fully in LDS:
//          inverse FFT along i2 on the active i1 columns only, inverse FFT along i1,
//          multiply by V, forward FFT along i1, forward FFT along i2 on the active
//          columns, keep the sphere columns.
//   zfwd   (grid ncol_pad/CT x nb): forward FFT along i3, keep the sphere entries,
//          scale 1/nnr, add kinetic * X, add the nonlocal result, apply
//          out = alpha*H + beta*X + gamma*Z.
// Transforms are Stockham autosort in LDS, radix 8/4/2/3/5, with each stage loading
// its butterflies into registers before a barrier so a single LDS buffer suffices.
#include <cstdlib>

#include "pwh_dft.h"

namespace pwh {

constexpr int pick_radix(int rem) {
  return (rem % 8 == 0 && rem != 16) ? 8
       : (rem % 4 == 0)              ? 4
       : (rem % 2 == 0)              ? 2
       : (rem % 3 == 0)              ? 3
       : (rem % 5 == 0)              ? 5
                                     : rem;
}

// One Stockham stage over `count` transforms of length N held in LDS.  Element e of
// transform t lives at buf[base(t) + e*es].  TFAST: consecutive threads take
// consecutive transforms (columns), else consecutive butterflies (rows).
template <typename T, int N, int R, int NS, int SIGN, int NT, int MAXCOUNT, bool TFAST,
          typename BaseF, typename Hooks>
__device__ __forceinline__ void lds_stage(cx<T>* buf, int count, int es, BaseF base,
                                          const cx<T>* tw, int tid, const Hooks& hk) {
  constexpr int NB = N / R;
  constexpr int MAXB = (MAXCOUNT * NB + NT - 1) / NT;
  constexpr bool FIRST = (NS == 1), LAST = (NS * R == N);
  const int total = count * NB;
  cx<T> v[MAXB][R];
  int off[MAXB], jv[MAXB], tv[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int idx = tid + q * NT;
    jv[q] = -1;
    off[q] = 0;
    tv[q] = 0;
    if (idx < total) {
      int t, j;
      if constexpr (TFAST) {
        j = idx / count;
        t = idx - j * count;
      } else {
        t = idx / NB;
        j = idx - t * NB;
      }
      const int b = base(t);
      off[q] = b;
      jv[q] = j;
      tv[q] = t;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (!FIRST || hk.live_in(t, j + r * NB)) v[q][r] = buf[b + (j + r * NB) * es];
        else v[q][r] = mk<T>(0, 0);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    if (jv[q] >= 0) {
      const int j = jv[q];
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = tw[k * r * (N / (NS * R))];
          if constexpr (SIGN < 0) w.y = -w.y;
          v[q][r] = v[q][r] * w;
        }
      }
      dft<T, R, SIGN>(v[q]);
      const int o = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = o + r * NS;
        if constexpr (LAST) {
          if (!hk.live_out(tv[q], e)) continue;
          hk.post(tv[q], e, v[q][r]);
        }
        buf[off[q] + e * es] = v[q][r];
      }
    }
  }
  __syncthreads();
}

template <typename T, int N, int SIGN, int NT, int MAXCOUNT, bool TFAST, int NS = 1, typename BaseF,
          typename Hooks = NoHooks>
__device__ __forceinline__ void lds_fft(cx<T>* buf, int count, int es, BaseF base, const cx<T>* tw,
                                        int tid, const Hooks& hk = Hooks()) {
  if constexpr (NS < N) {
    constexpr int R = pick_radix(N / NS);
    lds_stage<T, N, R, NS, SIGN, NT, MAXCOUNT, TFAST>(buf, count, es, base, tw, tid, hk);
    lds_fft<T, N, SIGN, NT, MAXCOUNT, TFAST, NS * R>(buf, count, es, base, tw, tid, hk);
  }
}

constexpr bool smooth235(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// ---------------------------------------------------------------------------
// Pass 1: inverse transform along i3 of the sphere columns
// ---------------------------------------------------------------------------
constexpr int kColTile = 16;      // column tile of the dense line transforms (k_fft_lines)
constexpr int kColThreads = 128;
constexpr int kPlaneThreads = 512;

// Column passes: a block transforms CT consecutive sphere columns of one band with NT
// lanes.  Every global load a block needs is issued before its first barrier: the tile's
// column range, the twiddles and up to PE sphere entries per lane (k_zinv), or the W tile
// and the epilogue operands (k_zfwd), so a block waits for one memory latency rather
// than a chain of them.  Entries beyond PE * NT in a tile (PE is picked from the largest
// tile) go through a plain loop.  The last tile may be short: cn = min(CT, ncol_pad - c0).
template <typename T, int NZ, int CT, int NT, int PE>
__device__ __forceinline__ void zinv_body(SphereMaps m, const cx<T>* __restrict__ tw3g,
                                             const cx<T>* __restrict__ X, int ldx,
                                             cx<T>* __restrict__ W, const int bx, const int by) {
  __shared__ cx<T> tile[NZ * CT];
  __shared__ cx<T> tw[NZ];
  constexpr int TWP = (NZ + NT - 1) / NT;
  const int tid = threadIdx.x;
  const int c0 = bx * CT;
  const int b = by;
  const int cn = min(CT, m.ncol_pad - c0);
  const int kb = m.colstart[c0], ke = m.colstart[c0 + cn];
  const cx<T>* xb = X + (size_t)b * ldx;
  // Loads use clamped, always valid indices and no branch, so they issue back to back;
  // out-of-range lanes are discarded afterwards.  Register copies are separate real
  // arrays (arrays of complex structs end up in scratch).
  T twr[TWP], twi[TWP];
#pragma unroll
  for (int j = 0; j < TWP; ++j) {
    const cx<T> v = tw3g[min(tid + j * NT, NZ - 1)];
    twr[j] = v.x; twi[j] = v.y;
  }
  const int klast = max(ke - 1, 0);
  T xr[PE], xi[PE];
  int kzv[PE], kcv[PE];
#pragma unroll
  for (int j = 0; j < PE; ++j) {
    const int k = min(kb + tid + j * NT, klast);
    const cx<T> v = xb[k];
    xr[j] = v.x; xi[j] = v.y;
    kzv[j] = m.kz[k];
    kcv[j] = m.kcol[k];
  }
  for (int i = tid; i < NZ * CT; i += NT) tile[i] = mk<T>(0, 0);
#pragma unroll
  for (int j = 0; j < TWP; ++j)
    if (tid + j * NT < NZ) tw[tid + j * NT] = mk<T>(twr[j], twi[j]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PE; ++j)
    if (kb + tid + j * NT < ke) tile[kzv[j] * CT + (kcv[j] - c0)] = mk<T>(xr[j], xi[j]);
  for (int k = kb + PE * NT + tid; k < ke; k += NT) tile[m.kz[k] * CT + (m.kcol[k] - c0)] = xb[k];
  __syncthreads();
  lds_fft<T, NZ, +1, NT, CT, true>(tile, CT, CT, [](int t) { return t; }, tw, tid);
  cx<T>* wb = W + (size_t)b * NZ * m.ncol_pad + c0;
  for (int i = tid; i < NZ * CT; i += NT) {
    const int z = i / CT, c = i - z * CT;
    if (c < cn) wb[(size_t)z * m.ncol_pad + c] = tile[i];
  }
}

template <typename T, int NZ, int CT, int NT, int PE>
__global__ __launch_bounds__(NT) void k_zinv(SphereMaps m, const cx<T>* __restrict__ tw3g,
                                             const cx<T>* __restrict__ X, int ldx,
                                             cx<T>* __restrict__ W) {
  zinv_body<T, NZ, CT, NT, PE>(m, tw3g, X, ldx, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Pass 2: (i1,i2) planes: inverse y, inverse x, *V, forward x, forward y
// ---------------------------------------------------------------------------
// The plane tile keeps rows (fixed i2) at an odd stride S1 = N1 + 1 complex: with
// transform-fast lane mapping a wave then touches 64 different rows at distinct banks
// in every stage of the row transforms, and consecutive columns in the column ones.
template <int N1>
constexpr int plane_stride() { return N1 + 1; }

template <typename T, int N1, int N2>
constexpr size_t plane_smem() {
  return sizeof(cx<T>) * (size_t)(N2 * plane_stride<N1>() + N1 + N2) + sizeof(int) * N1 + N1;
}

template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void plane_body(SphereMaps m, const cx<T>* __restrict__ tw1g,
                                              const cx<T>* __restrict__ tw2g,
                                              const T* __restrict__ V, cx<T>* __restrict__ W, const int bx, const int by) {
  constexpr int S1 = plane_stride<N1>();
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw1 = tile + N2 * S1;
  cx<T>* tw2 = tw1 + N1;
  int* xa = reinterpret_cast<int*>(tw2 + N2);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int z = bx, b = by;
  const int nxa = m.nxa;
  for (int i = tid; i < N1; i += NT) { tw1[i] = tw1g[i]; xact[i] = 0; }
  for (int i = tid; i < N2; i += NT) tw2[i] = tw2g[i];
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  // only the active columns need zeros: the x transforms never read the others
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
  lds_fft<T, N2, +1, NT, N1, true>(tile, nxa, S1, [&](int t) { return xa[t]; }, tw2, tid);
  const XInvHooks<T> hinv{xact, V + (size_t)z * N1 * N2, N1};
  lds_fft<T, N1, +1, NT, N2, true>(tile, N2, 1, [](int t) { return t * S1; }, tw1, tid, hinv);
  const XFwdHooks hfwd{xact};
  lds_fft<T, N1, -1, NT, N2, true>(tile, N2, 1, [](int t) { return t * S1; }, tw1, tid, hfwd);
  lds_fft<T, N2, -1, NT, N1, true>(tile, nxa, S1, [&](int t) { return xa[t]; }, tw2, tid);
  for (int c = tid; c < m.ncol; c += NT) {
    const int xy = m.colxy[c];
    wrow[c] = tile[(xy / N1) * S1 + xy % N1];
  }
}

template <typename T, int N1, int N2, int NT>
__global__ __launch_bounds__(NT) void k_plane(SphereMaps m, const cx<T>* __restrict__ tw1g,
                                              const cx<T>* __restrict__ tw2g,
                                              const T* __restrict__ V, cx<T>* __restrict__ W) {
  plane_body<T, N1, N2, NT>(m, tw1g, tw2g, V, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Pass 2, band-pruned: the sphere's i1 and i2 lie in the band [-n/4, n/4)
// ---------------------------------------------------------------------------
// With a first inverse radix R in {4, 8} (n % 4 == 0), butterfly j of the first stage
// reads elements j + r n/R, which lie in the band exactly for r < R/4 or r >= R - R/4:
// half of its inputs are known zeros.  The forward transforms run the radices in
// reverse order, so their last stage has the same R and half of its outputs (those
// outside the band) are never needed.  The column transforms run over the n/2 band
// columns (a compile-time count) instead of the active ones.

constexpr int nstages(int n) {
  int s = 0;
  for (int rem = n; rem > 1; rem /= pick_radix(rem)) ++s;
  return s;
}
constexpr int radix_nat(int n, int i) {
  int rem = n;
  for (int s = 0; s < i; ++s) rem /= pick_radix(rem);
  return pick_radix(rem);
}
constexpr int radix_at(int n, int i, bool rev) { return rev ? radix_nat(n, nstages(n) - 1 - i) : radix_nat(n, i); }
constexpr int ns_at(int n, int i, bool rev) {
  int ns = 1;
  for (int s = 0; s < i; ++s) ns *= radix_at(n, s, rev);
  return ns;
}
constexpr bool band_live(int R, int r) { return r < R / 4 || r >= R - R / 4; }

// DFT of 4 with v[1] = v[2] = 0
template <typename T, int SIGN>
__device__ __forceinline__ void dft4_in03(cx<T>* v) {
  const cx<T> a = v[0], d = v[3], sd = mul_si<SIGN>(d);
  v[0] = a + d; v[1] = a - sd; v[2] = a - d; v[3] = a + sd;
}
// DFT of 8 with v[2..5] = 0 (even/odd split, each half a two-input DFT of 4)
template <typename T, int SIGN>
__device__ __forceinline__ void dft8_in0167(cx<T>* v) {
  const cx<T> e0 = v[0], e3 = v[6], o0 = v[1], o3 = v[7];
  const cx<T> se = mul_si<SIGN>(e3), so = mul_si<SIGN>(o3);
  const cx<T> E0 = e0 + e3, E1 = e0 - se, E2 = e0 - e3, E3 = e0 + se;
  const cx<T> O0 = o0 + o3, O1 = o0 - so, O2 = o0 - o3, O3 = o0 + so;
  const T c = T(0.70710678118654752440084436210485);
  const cx<T> w1 = mk<T>(c * (O1.x - SIGN * O1.y), c * (O1.y + SIGN * O1.x));
  const cx<T> w2 = mul_si<SIGN>(O2);
  const cx<T> w3 = mk<T>(c * (-O3.x - SIGN * O3.y), c * (-O3.y + SIGN * O3.x));
  v[0] = E0 + O0; v[4] = E0 - O0;
  v[1] = E1 + w1; v[5] = E1 - w1;
  v[2] = E2 + w2; v[6] = E2 - w2;
  v[3] = E3 + w3; v[7] = E3 - w3;
}
// DFT of 4, outputs 0 and 3 only
template <typename T, int SIGN>
__device__ __forceinline__ void dft4_out03(cx<T>* v) {
  const cx<T> a = v[0] + v[2], b = v[1] + v[3], c = v[0] - v[2], d = mul_si<SIGN>(v[1] - v[3]);
  v[0] = a + b; v[3] = c - d;
}
// DFT of 8, outputs 0, 1, 6, 7 only: X[2m] = DFT4(v_n + v_n+4)[m], X[2m+1] =
// DFT4((v_n - v_n+4) w^n)[m] for m in {0, 3}
template <typename T, int SIGN>
__device__ __forceinline__ void dft8_out0167(cx<T>* v) {
  cx<T> s[4], d[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) { s[n] = v[n] + v[n + 4]; d[n] = v[n] - v[n + 4]; }
  const T c = T(0.70710678118654752440084436210485);
  d[1] = mk<T>(c * (d[1].x - SIGN * d[1].y), c * (d[1].y + SIGN * d[1].x));
  d[2] = mul_si<SIGN>(d[2]);
  d[3] = mk<T>(c * (-d[3].x - SIGN * d[3].y), c * (-d[3].y + SIGN * d[3].x));
  dft4_out03<T, SIGN>(s);
  dft4_out03<T, SIGN>(d);
  v[0] = s[0]; v[6] = s[3]; v[1] = d[0]; v[7] = d[3];
}

struct NoPost {
  template <typename V>
  __device__ __forceinline__ void operator()(int, int, V&) const {}
};

// One Stockham stage (radix R, NS = product of the earlier radices) over COUNT transforms
// of length N; element e of transform t lives at buf[base(t) + e*ES].  Consecutive lanes
// take consecutive transforms.  PR: 0 full, 1 inputs outside the band are zero, 2 only
// outputs in the band are stored.  When COUNT is a multiple of 64 every wave works on
// one butterfly index j, so its twiddles are read as scalars from twg.
template <typename T, int N, int R, int NS, int SIGN, int NT, int COUNT, int ES, int PR, bool LASTST,
          typename BaseF, typename Post>
__device__ __forceinline__ void band_stage(cx<T>* buf, BaseF base, const cx<T>* tw,
                                           const cx<T>* __restrict__ twg, int tid, const Post& post) {
  constexpr int NB = N / R;
  constexpr int TOTAL = COUNT * NB;
  constexpr int MAXB = (TOTAL + NT - 1) / NT;
  constexpr bool UNIFORM = (COUNT % 64 == 0);
  static_assert(PR == 0 || R == 4 || R == 8, "band pruning needs a radix-4 or radix-8 stage");
  cx<T> v[MAXB][R];
  int tb[MAXB], jv[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int idx = tid + q * NT;
    jv[q] = -1;
    tb[q] = 0;
    if (TOTAL % NT == 0 || idx < TOTAL) {
      int j = idx / COUNT;
      if constexpr (UNIFORM) j = __builtin_amdgcn_readfirstlane(j);
      const int t = idx - j * COUNT;
      const int b = base(t);
      tb[q] = b;
      jv[q] = j;
      const int bo = b + j * ES;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (PR == 1 && !band_live(R, r)) v[q][r] = mk<T>(0, 0);
        else v[q][r] = buf[bo + r * NB * ES];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    if (jv[q] >= 0) {
      int j = jv[q];
      if constexpr (UNIFORM) j = __builtin_amdgcn_readfirstlane(j);
      const int k = j % NS;
      if constexpr (NS > 1) {
        const int ks = k * (N / (NS * R));
        int ti = 0;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          ti += ks;
          cx<T> w = UNIFORM ? twg[ti] : tw[ti];
          if constexpr (SIGN < 0) w.y = -w.y;
          v[q][r] = v[q][r] * w;
        }
      }
      if constexpr (PR == 1 && R == 8) dft8_in0167<T, SIGN>(v[q]);
      else if constexpr (PR == 1 && R == 4) dft4_in03<T, SIGN>(v[q]);
      else if constexpr (PR == 2 && R == 8) dft8_out0167<T, SIGN>(v[q]);
      else if constexpr (PR == 2 && R == 4) dft4_out03<T, SIGN>(v[q]);
      else dft<T, R, SIGN>(v[q]);
      const int o = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (PR == 2 && !band_live(R, r)) continue;
        if constexpr (LASTST) post(q, r, v[q][r]);
        buf[tb[q] + (o + r * NS) * ES] = v[q][r];
      }
    }
  }
  __syncthreads();
}

// Full transform: stages I..nstages(N)-1.  PIN prunes the first stage's inputs, POUT the
// last stage's outputs; REV runs the radices in reverse order.
template <typename T, int N, int SIGN, bool REV, int NT, int COUNT, int ES, bool PIN, bool POUT,
          int I = 0, typename BaseF, typename Post>
__device__ __forceinline__ void band_fft(cx<T>* buf, BaseF base, const cx<T>* tw, const cx<T>* __restrict__ twg,
                                         int tid, const Post& post) {
  constexpr int NSTG = nstages(N);
  if constexpr (I < NSTG) {
    constexpr int R = radix_at(N, I, REV), NS = ns_at(N, I, REV);
    constexpr int PR = (PIN && I == 0) ? 1 : (POUT && I == NSTG - 1) ? 2 : 0;
    band_stage<T, N, R, NS, SIGN, NT, COUNT, ES, PR, I == NSTG - 1>(buf, base, tw, twg, tid, post);
    band_fft<T, N, SIGN, REV, NT, COUNT, ES, PIN, POUT, I + 1>(buf, base, tw, twg, tid, post);
  }
}

template <int N>
constexpr int band_plane_threads() {
  const int w = N * (N / radix_nat(N, 0));
  const int r = (w + 63) / 64 * 64;
  return r < 512 ? r : 512;
}

template <typename T, int N>
constexpr size_t band_plane_smem() {
  return sizeof(cx<T>) * ((size_t)N * (N + 1) + N);
}

template <typename T, int N, int NT>
__device__ __forceinline__ void plane_band_body(SphereMaps m, const cx<T>* __restrict__ twg,
                                                   const T* __restrict__ V, cx<T>* __restrict__ W, const int bx, const int by) {
  constexpr int S1 = N + 1, Q = N / 4, NBC = N / 2;
  static_assert(N % 4 == 0, "band pruning needs n % 4 == 0");
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  cx<T>* tw = tile + N * S1;
  const int tid = threadIdx.x;
  const int z = bx, b = by;

  // V for the outputs this lane produces in the last inverse row stage, loaded first so
  // the latency hides behind the column transforms
  constexpr int RL = radix_at(N, nstages(N) - 1, false), NSL = N / RL;
  constexpr int TOTL = N * NSL, MBL = (TOTL + NT - 1) / NT;
  T vp[MBL][RL];
  const T* Vz = V + (size_t)z * N * N;
#pragma unroll
  for (int q = 0; q < MBL; ++q) {
    const int idx = tid + q * NT;
    if (TOTL % NT == 0 || idx < TOTL) {
      const int j = idx / N, t = idx - j * N;
#pragma unroll
      for (int r = 0; r < RL; ++r) vp[q][r] = Vz[t * N + j + r * NSL];
    }
  }

  auto band_x = [](int t) { return t < Q ? t : t + NBC; };
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  for (int i = tid; i < NBC * NBC; i += NT) {
    const int yy = i / NBC, xx = i - yy * NBC;
    tile[band_x(yy) * S1 + band_x(xx)] = mk<T>(0, 0);
  }
  __syncthreads();
  cx<T>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  for (int c = tid; c < m.ncol; c += NT) {
    const int xy = m.colxy[c];
    tile[(xy / N) * S1 + xy % N] = wrow[c];
  }
  __syncthreads();
  const NoPost none;
  auto col_base = [&](int t) { return band_x(t); };
  auto row_base = [](int t) { return t * S1; };
  auto vmul = [&](int q, int r, cx<T>& v) { v = vp[q][r] * v; };
  // inverse along i2 on the band columns, inverse along i1 on every row (times V),
  // forward along i1 keeping band columns, forward along i2 keeping band rows
  band_fft<T, N, +1, false, NT, NBC, S1, true, false>(tile, col_base, tw, twg, tid, none);
  band_fft<T, N, +1, false, NT, N, 1, true, false>(tile, row_base, tw, twg, tid, vmul);
  band_fft<T, N, -1, true, NT, N, 1, false, true>(tile, row_base, tw, twg, tid, none);
  band_fft<T, N, -1, true, NT, NBC, S1, false, true>(tile, col_base, tw, twg, tid, none);
  for (int c = tid; c < m.ncol; c += NT) {
    const int xy = m.colxy[c];
    wrow[c] = tile[(xy / N) * S1 + xy % N];
  }
}

template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void k_plane_band(SphereMaps m, const cx<T>* __restrict__ twg,
                                                   const T* __restrict__ V, cx<T>* __restrict__ W) {
  plane_band_body<T, N, NT>(m, twg, V, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Pass 2 for 64 x 64 planes with the sphere in the band: wave-local four-step transforms
// ---------------------------------------------------------------------------
// Each 64-point line is an 8 x 8 four-step FFT owned by 8 lanes of one wave.  With
// x[j + 8r] in lane j (register r):  y_j[s] = w64^(js) DFT8_r(x[j + 8r])[s], then
// X[s + 8t] = DFT8_j(y_j[s])[t] in lane s.  The transpose between the two DFT8 steps goes
// through LDS within the wave (wave-level ordering, no workgroup barrier), and lane s
// ends up holding X[s + 8t] -- exactly the inputs of the next line transform in the same
// direction, so inverse i1, x V and forward i1 chain in registers.
//
// The column passes read the plane's sphere columns from W and write them back through
// the band map (band row, band column) -> column index, so the plane needs no zero fill,
// scatter or gather phase and two workgroup barriers in all.
//
// Layout: LDS serves 16 consecutive lanes of 8-byte words (8 of 16-byte words) per cycle,
// so consecutive lanes must fall on distinct banks.  The tile has 64 rows of 72 complex
// (8 mod 16).  Every LDS address below is a per-lane base plus a compile-time offset:
//   - column passes, waves 0-3 (lane l: band column 8 wave + l % 8, j = l / 8): the data at
//     rows j + 8r of the band column; the exchange in the wave's own rows
//     [16 wave, 16 wave + 16) of the non-band columns 16..47, y_j[s] at row
//     16 wave + 8 (s >> 2) + j, column 16 + l % 8 + 8 (s & 3);
//   - row pass, all 8 waves (lane l: j = l % 8, row 8 wave + l / 8): the row itself, band
//     entries at j + 8r, y_j[s] at slot 9s + j (9 is odd and a row is 8 mod 16, so the
//     two rows of 8 lanes in a group of 16 hit distinct banks).
constexpr int kW64Stride = 72;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exchange offsets: lane j writes y_j[s] at wr[w64_woff(s)], lane s reads y_i[s] at
// rd[w64_roff(i)].
template <bool ROW>
__device__ __forceinline__ constexpr int w64_woff(int s) {
  return ROW ? 9 * s : 8 * (s >> 2) * kW64Stride + 8 * (s & 3);
}
template <bool ROW>
__device__ __forceinline__ constexpr int w64_roff(int i) { return ROW ? i : i * kW64Stride; }

// The four-step line transform on 8 lanes: lane j passes x[j + 8r] in v (only r in
// {0, 1, 6, 7} when PIN: the band inputs), gets X[j + 8t] back in v (only t in
// {0, 1, 6, 7} when POUT).  tw[r] = exp(+2 pi i j r / 64) (r = 1..7) for this lane;
// SIGN < 0 uses its conjugate.
template <typename T, int SIGN, bool PIN, bool POUT, bool ROW>
__device__ __forceinline__ void w64_line(cx<T> (&v)[8], cx<T>* wr, const cx<T>* rd, const cx<T> (&tw)[8]) {
  if constexpr (PIN) dft8_in0167<T, SIGN>(v);
  else dft<T, 8, SIGN>(v);
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = v[r] * (SIGN > 0 ? tw[r] : conjg(tw[r]));
#pragma unroll
  for (int r = 0; r < 8; ++r) wr[w64_woff<ROW>(r)] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = rd[w64_roff<ROW>(i)];
  if constexpr (POUT) dft8_out0167<T, SIGN>(v);
  else dft<T, 8, SIGN>(v);
}

// twg: the length-64 twiddle table exp(+2 pi i k / 64), k < 64, followed by the 8 x 8
// table [j][r] = exp(+2 pi i j r / 64).  bandmap[yb * 32 + xb]: the column index of band
// row yb and band column xb (band index b is position b for b < 16, b + 32 above), or -1.
template <typename T>
__device__ __forceinline__ void plane_w64_body(SphereMaps m, const cx<T>* __restrict__ twg,
                                                   const T* __restrict__ V, cx<T>* __restrict__ W, const int bx, const int by) {
  constexpr int N = 64, S = kW64Stride;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int z = bx, b = by;
  const int jc = l >> 3, cl = l & 7;   // column passes: band column 8 wave + cl, j = jc
  const int jr = l & 7, yl = l >> 3;   // row pass: row 8 wave + yl, j = jr
  // V at the row pass's points (row 8 wave + yl, x = jr + 8t), loaded first
  const int y = 8 * wave + yl;
  T vr[8];
  const T* Vrow = V + ((size_t)z * N + y) * N;
#pragma unroll
  for (int t = 0; t < 8; ++t) vr[t] = Vrow[jr + 8 * t];
  const cx<T>* tw8 = twg + N;
  cx<T> twi[8];
  auto load_tw = [&](int j) {
    twi[0] = mk<T>(1, 0);
#pragma unroll
    for (int r = 1; r < 8; ++r) twi[r] = tw8[8 * j + r];
  };
  cx<T>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  const int xb = 8 * (wave & 3) + cl;
  const int* bm = m.bandmap + jc * 32 + xb;   // band rows jc + 8q, q = 0..3: + 256 q
  cx<T>* col = tile + (xb < 16 ? xb : xb + 32);
  cx<T>* xwr = tile + (16 * wave + jc) * S + 16 + cl;
  const cx<T>* xrd = tile + (16 * wave + 8 * (jc >> 2)) * S + 16 + cl + 8 * (jc & 3);
  cx<T> v[8];
  // inverse along i2 on the 32 band columns (inputs in the band rows 0, 1, 6, 7 of 8)
  if (wave < 4) {
    int c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = bm[256 * q];
    load_tw(jc);
    cx<T> e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) e[q] = wrow[c[q] < 0 ? 0 : c[q]];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q < 2 ? q : q + 4] = c[q] < 0 ? mk<T>(0, 0) : e[q];
    v[2] = v[3] = v[4] = v[5] = mk<T>(0, 0);
    w64_line<T, +1, true, false, false>(v, xwr, xrd, twi);
#pragma unroll
    for (int t = 0; t < 8; ++t) col[(jc + 8 * t) * S] = v[t];
  }
  __syncthreads();
  // rows: inverse along i1 (inputs in the band columns), times V, forward along i1 keeping
  // the band columns
  {
    load_tw(jr);
    cx<T>* row = tile + y * S;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = (r < 2 || r >= 6) ? row[jr + 8 * r] : mk<T>(0, 0);
    w64_line<T, +1, true, false, true>(v, row + jr, row + 9 * jr, twi);
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = vr[t] * v[t];
    w64_line<T, -1, false, true, true>(v, row + jr, row + 9 * jr, twi);
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (t < 2 || t >= 6) row[jr + 8 * t] = v[t];
  }
  __syncthreads();
  // forward along i2 on the band columns, keeping the band rows, written to W
  if (wave < 4) {
    int c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = bm[256 * q];
    load_tw(jc);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = col[(jc + 8 * r) * S];
    w64_line<T, -1, false, true, false>(v, xwr, xrd, twi);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c[q] >= 0) wrow[c[q]] = v[q < 2 ? q : q + 4];
  }
}

template <typename T>
__global__ __launch_bounds__(512) void k_plane_w64(SphereMaps m, const cx<T>* __restrict__ twg,
                                                   const T* __restrict__ V, cx<T>* __restrict__ W) {
  plane_w64_body<T>(m, twg, V, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Pass 2 for 96 x 96 planes with the sphere in the band: 8-lane x 12-value line transforms
// ---------------------------------------------------------------------------
// A 96-point line is owned by 8 lanes of one wave, 12 values per lane.  With x[j + 8r] in
// lane j (register r), the inverse transform X[s + 12t] = sum_j w8^(jt) w96^(js) A_j[s],
// A_j = DFT12_r(x[j + 8r]), runs as three steps:
//   1. DFT-12 over r in registers, times w96^(js);
//   2. radix 2 over a (j = b + 4a): lane (h, b) = (l / 4, l % 4) takes s in [6h, 6h + 6)
//      of lanes b and b + 4, and forms C_c = B_b + (-1)^c B_(b+4) times w8^(bc);
//   3. radix 4 over b: lane (h, b') takes the (s, c) pairs q = 2(s - 6h) + c in
//      [3b', 3b' + 3) of lanes (h, 0..3) and forms X[s + 12c + 24d], d = 0..3.
// The two exchanges go through LDS within the wave.  Lane l ends with X[k] for
// k = w96_k(l, u, d) in register 4u + d, which is the input layout of the forward
// transform (the same steps in reverse), so inverse i1, x V and forward i1 chain in
// registers.  Exchange slots are a per-lane base plus compile-time offsets:
//   step 1 -> 2, element B_j[s]: j + 8 (s % 6) + H1 (s / 6);
//   step 2 -> 3, element (h, b, q): 4h + b + 8 (q % 3) + B2 (q / 3).
// Rows (H1 = 52, B2 = 25, 100 slots of a row of 104) are conflict-free: the 8 lanes of a
// row hit 8 different banks mod 8 in every access, and the row stride (8 mod 16) puts the
// second row of a 16-lane group on the other half.  Columns exchange within the column's
// own 96 slots (H1 = 48, B2 = 24), the band columns of a wave are consecutive, and the
// few accesses where two line lanes of a 16-lane group land on rows of equal parity
// conflict two-way.
constexpr int kW96Stride = 104;

__device__ __forceinline__ int w96_k(int l, int u, int d) {
  const int q = 3 * (l & 3) + u;
  return 6 * (l >> 2) + (q >> 1) + 12 * (q & 1) + 24 * d;
}

template <typename T>
__device__ __forceinline__ cx<T> cmul_k(cx<T> a, T wr, T wi) { return mk<T>(a.x * wr - a.y * wi, a.x * wi + a.y * wr); }

// DFT of 12 as 3 x 4: r = 4a + b, k = c + 3d, X[c + 3d] = sum_b w4^(bd) w12^(bc) DFT3_a(x[4a + b])[c].
// PIN: x[3..8] = 0 (the band inputs 0, 1, 2, 9, 10, 11); POUT: only X[0, 1, 2, 9, 10, 11].
template <typename T, int SIGN, bool PIN, bool POUT>
__device__ __forceinline__ void dft12(cx<T> (&v)[12]) {
  const T h = T(0.86602540378443864676372317075294);
  cx<T> y[4][3];
  if constexpr (PIN) {
    // x * w3^2 and x * w3, w3 = -1/2 + SIGN i sqrt(3)/2
    auto w3sq = [&](cx<T> a) { return mk<T>(T(-0.5) * a.x + SIGN * h * a.y, T(-0.5) * a.y - SIGN * h * a.x); };
    auto w3m = [&](cx<T> a) { return mk<T>(T(-0.5) * a.x - SIGN * h * a.y, T(-0.5) * a.y + SIGN * h * a.x); };
    y[0][0] = y[0][1] = y[0][2] = v[0];
    y[1][0] = v[1] + v[9]; y[1][1] = v[1] + w3sq(v[9]); y[1][2] = v[1] + w3m(v[9]);
    y[2][0] = v[2] + v[10]; y[2][1] = v[2] + w3sq(v[10]); y[2][2] = v[2] + w3m(v[10]);
    y[3][0] = v[11]; y[3][1] = w3sq(v[11]); y[3][2] = w3m(v[11]);
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      cx<T> t[3] = {v[b], v[4 + b], v[8 + b]};
      dft<T, 3, SIGN>(t);
      y[b][0] = t[0]; y[b][1] = t[1]; y[b][2] = t[2];
    }
  }
  // w12^(bc): w12 = (h, SIGN/2), w12^2 = (1/2, SIGN h), w12^3 = SIGN i, w12^4 = (-1/2, SIGN h), w12^6 = -1
  y[1][1] = cmul_k<T>(y[1][1], h, SIGN * T(0.5));
  y[1][2] = cmul_k<T>(y[1][2], T(0.5), SIGN * h);
  y[2][1] = cmul_k<T>(y[2][1], T(0.5), SIGN * h);
  y[2][2] = cmul_k<T>(y[2][2], T(-0.5), SIGN * h);
  y[3][1] = mul_si<SIGN>(y[3][1]);
  y[3][2] = mk<T>(-y[3][2].x, -y[3][2].y);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cx<T> t[4] = {y[0][c], y[1][c], y[2][c], y[3][c]};
    if constexpr (POUT) {
      dft4_out03<T, SIGN>(t);
      v[c] = t[0]; v[c + 9] = t[3];
    } else {
      dft<T, 4, SIGN>(t);
#pragma unroll
      for (int d = 0; d < 4; ++d) v[c + 3 * d] = t[d];
    }
  }
}

template <bool ROW> __device__ __forceinline__ constexpr int w96_h1() { return ROW ? 52 : 48; }
template <bool ROW> __device__ __forceinline__ constexpr int w96_b2() { return ROW ? 25 : 24; }

// c ? a : b component by component (a select of the whole struct can go through scratch
// memory for FP64)
template <typename T>
__device__ __forceinline__ cx<T> sel_cx(bool c, cx<T> a, cx<T> b) { return mk<T>(c ? a.x : b.x, c ? a.y : b.y); }

// w8^b (SIGN) for b = 0..3 at compile time
template <typename T, int SIGN, int B>
__device__ __forceinline__ cx<T> w8_mul(cx<T> a) {
  const T c = T(0.70710678118654752440084436210485);
  if constexpr (B == 0) return a;
  else if constexpr (B == 1) return cmul_k<T>(a, c, SIGN * c);
  else if constexpr (B == 2) return mul_si<SIGN>(a);
  else return cmul_k<T>(a, -c, SIGN * c);
}

// Inverse (SIGN > 0) line transform: lane l passes x[l + 8r] in v[r] (PIN: only r in
// {0, 1, 2, 9, 10, 11}); gets X[w96_k(l, u, d)] back in v[4u + d].  line: the line's
// exchange storage (element e at line[e * ES]); tw96[12 j + s] = w96^(js).
template <typename T, int SIGN, bool PIN, bool ROW, int CS = kW96Stride>
__device__ __forceinline__ void w96_inv(cx<T> (&v)[12], cx<T>* line, int l, const cx<T>* __restrict__ tw96,
                                        const cx<T>* __restrict__ twg) {
  constexpr int ES = ROW ? 1 : CS, H1 = w96_h1<ROW>(), B2 = w96_b2<ROW>();
  // every phase that hands data between lanes is fenced: LDS serves a wave's instructions
  // in order, and the fence keeps the compiler from moving a lane's write above another
  // lane's earlier read of the same slot
  wave_lds_sync();
  dft12<T, SIGN, PIN, false>(v);
  const cx<T>* tl = tw96 + 12 * l;
#pragma unroll
  for (int s = 1; s < 12; ++s) v[s] = v[s] * (SIGN > 0 ? tl[s] : conjg(tl[s]));
  {
    cx<T>* wr = line + l * ES;
#pragma unroll
    for (int s = 0; s < 12; ++s) wr[(8 * (s % 6) + H1 * (s / 6)) * ES] = v[s];
  }
  wave_lds_sync();
  const int b = l & 3, hh = l >> 2;
  const cx<T> w8b = SIGN > 0 ? twg[12 * b] : conjg(twg[12 * b]);
  {
    const cx<T>* rd = line + (b + H1 * hh) * ES;
#pragma unroll
    for (int sg = 0; sg < 6; ++sg) {
      const cx<T> x0 = rd[(8 * sg) * ES], x1 = rd[(4 + 8 * sg) * ES];
      v[2 * sg] = x0 + x1;
      v[2 * sg + 1] = (x0 - x1) * w8b;
    }
  }
  wave_lds_sync();
  {
    cx<T>* wr = line + (4 * hh + b) * ES;
#pragma unroll
    for (int q = 0; q < 12; ++q) wr[(8 * (q % 3) + B2 * (q / 3)) * ES] = v[q];
  }
  wave_lds_sync();
  {
    const cx<T>* rd = line + (4 * hh + B2 * b) * ES;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      cx<T> t[4];
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) t[bb] = rd[(bb + 8 * u) * ES];
      dft<T, 4, SIGN>(t);
#pragma unroll
      for (int d = 0; d < 4; ++d) v[4 * u + d] = t[d];
    }
  }
}

// Forward (SIGN < 0) line transform, the steps of w96_inv in reverse: lane l passes
// Y[w96_k(l, u, d)] in v[4u + d]; gets F[l + 8r] back in v[r] (POUT: only r in
// {0, 1, 2, 9, 10, 11} are formed).
template <typename T, int SIGN, bool POUT, bool ROW, int CS = kW96Stride>
__device__ __forceinline__ void w96_fwd(cx<T> (&v)[12], cx<T>* line, int l, const cx<T>* __restrict__ tw96) {
  constexpr int ES = ROW ? 1 : CS, H1 = w96_h1<ROW>(), B2 = w96_b2<ROW>();
  const int b = l & 3, hh = l >> 2;
  wave_lds_sync();
  {
    // radix 4 over d for each (s, c) pair q = 3b + u, then w8^(bb c) (c = q % 2)
    cx<T>* wr = line + (4 * hh + B2 * b) * ES;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      cx<T> t[4] = {v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]};
      dft<T, 4, SIGN>(t);
      const bool c1 = ((b + u) & 1) != 0;
      wr[(0 + 8 * u) * ES] = t[0];
      const cx<T> t1 = w8_mul<T, SIGN, 1>(t[1]), t2 = w8_mul<T, SIGN, 2>(t[2]), t3 = w8_mul<T, SIGN, 3>(t[3]);
      wr[(1 + 8 * u) * ES] = sel_cx(c1, t1, t[1]);
      wr[(2 + 8 * u) * ES] = sel_cx(c1, t2, t[2]);
      wr[(3 + 8 * u) * ES] = sel_cx(c1, t3, t[3]);
    }
  }
  wave_lds_sync();
  {
    // radix 2 over c: U_(b + 4a)[6h + sg] = z0 + (-1)^a z1
    const cx<T>* rd = line + (4 * hh + b) * ES;
#pragma unroll
    for (int sg = 0; sg < 6; ++sg) {
      const int q0 = 2 * sg, q1 = 2 * sg + 1;
      const cx<T> z0 = rd[(8 * (q0 % 3) + B2 * (q0 / 3)) * ES], z1 = rd[(8 * (q1 % 3) + B2 * (q1 / 3)) * ES];
      v[sg] = z0 + z1;
      v[6 + sg] = z0 - z1;
    }
  }
  wave_lds_sync();
  {
    cx<T>* wr = line + (b + H1 * hh) * ES;
#pragma unroll
    for (int sg = 0; sg < 6; ++sg) {
      wr[(8 * sg) * ES] = v[sg];
      wr[(4 + 8 * sg) * ES] = v[6 + sg];
    }
  }
  wave_lds_sync();
  {
    const cx<T>* rd = line + l * ES;
#pragma unroll
    for (int s = 0; s < 12; ++s) v[s] = rd[(8 * (s % 6) + H1 * (s / 6)) * ES];
  }
  const cx<T>* tl = tw96 + 12 * l;
#pragma unroll
  for (int s = 1; s < 12; ++s) v[s] = v[s] * (SIGN > 0 ? tl[s] : conjg(tl[s]));
  dft12<T, SIGN, false, POUT>(v);
}

// twg: the length-96 twiddle table exp(+2 pi i k / 96), k < 96, followed by the 8 x 12
// table [j][s] = exp(+2 pi i j s / 96).  bandmap[yb * 48 + xb]: the column index of band
// row yb and band column xb (band index b is position b for b < 24, b + 48 above), or -1.
// Waves 0-5 run the column passes (band column 8 wave + l % 8, line lane l / 8), all 12
// waves the row pass (row 8 wave + l / 8, line lane l % 8).
template <typename T>
__device__ __forceinline__ void plane_w96_body(SphereMaps m, const cx<T>* __restrict__ twg,
                                               const T* __restrict__ V, cx<T>* __restrict__ W, const int bx,
                                               const int by) {
  constexpr int N = 96, S = kW96Stride, NB = 48;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int z = bx, b = by;
  const cx<T>* tw96 = twg + N;
  const int jr = l & 7, y = 8 * wave + (l >> 3);
  // V at the row pass's inverse outputs, loaded first
  T vr[12];
  const T* Vrow = V + ((size_t)z * N + y) * N;
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int d = 0; d < 4; ++d) vr[4 * u + d] = Vrow[w96_k(jr, u, d)];
  cx<T>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  const int jc = l >> 3, cl = l & 7;
  const int xb = 8 * (wave < 6 ? wave : 0) + cl;
  cx<T>* col = tile + (xb < 24 ? xb : xb + 48);
  const int* bm = m.bandmap + jc * NB + xb;   // band rows jc + 8i, i = 0..5: + 8 NB i
  cx<T> v[12];
  // inverse along i2 on the 48 band columns (inputs in the band rows), into the tile
  if (wave < 6) {
    int c[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = bm[8 * NB * i];
    cx<T> e[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) e[i] = wrow[c[i] < 0 ? 0 : c[i]];
#pragma unroll
    for (int r = 0; r < 12; ++r) v[r] = mk<T>(0, 0);
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i < 3 ? i : i + 6] = c[i] < 0 ? mk<T>(0, 0) : e[i];
    w96_inv<T, +1, true, false>(v, col, jc, tw96, twg);
    wave_lds_sync();
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int d = 0; d < 4; ++d) col[w96_k(jc, u, d) * S] = v[4 * u + d];
  }
  __syncthreads();
  // rows: inverse along i1 (inputs in the band columns), times V, forward along i1 keeping
  // the band columns
  {
    cx<T>* row = tile + y * S;
#pragma unroll
    for (int r = 0; r < 12; ++r) v[r] = (r < 3 || r >= 9) ? row[jr + 8 * r] : mk<T>(0, 0);
    w96_inv<T, +1, true, true>(v, row, jr, tw96, twg);
#pragma unroll
    for (int i = 0; i < 12; ++i) v[i] = vr[i] * v[i];
    w96_fwd<T, -1, true, true>(v, row, jr, tw96);
#pragma unroll
    for (int r = 0; r < 12; ++r)
      if (r < 3 || r >= 9) row[jr + 8 * r] = v[r];
  }
  __syncthreads();
  // forward along i2 on the band columns, keeping the band rows, written to W
  if (wave < 6) {
    int c[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = bm[8 * NB * i];
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int d = 0; d < 4; ++d) v[4 * u + d] = col[w96_k(jc, u, d) * S];
    w96_fwd<T, -1, true, false>(v, col, jc, tw96);
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if (c[i] >= 0) wrow[c[i]] = v[i < 3 ? i : i + 6];
  }
}

template <typename T>
__global__ __launch_bounds__(768) void k_plane_w96(SphereMaps m, const cx<T>* __restrict__ twg,
                                                   const T* __restrict__ V, cx<T>* __restrict__ W) {
  plane_w96_body<T>(m, twg, V, W, blockIdx.x, blockIdx.y);
}

// FP64 96 x 96 planes.  An FP64 plane with full rows would fill a CU's LDS, so the tile
// holds only the 48 band columns of each row (52 complex = 104 eight-byte words per row,
// 78 KB per plane: two planes per CU), and the row transforms exchange the real parts and
// then the imaginary parts through 100 of the row's words with the row layouts above
// (the row stride, 104 words, is 8 mod 16, as the FP32 rows').  Columns exchange complex
// values in place, as in FP32.  A workgroup is 6 waves: the 48 band columns, then the 96
// rows in two rounds of 48.
constexpr int kW96dStride = 52;

// The value unchanged, but opaque to the optimiser: per-lane table reads made through it
// are not merged with the same reads of an earlier pass, whose results would otherwise be
// kept live (and spilled) across the passes in between.
__device__ __forceinline__ int w96d_opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// The 12-value exchange of one row in two halves, in place: lane writes part of v[i] at
// wr[WO(i)], then reads part of v[i] from rd[RO(i)]; the real parts first (the imaginary
// parts are still the inputs when they are written).
template <typename WO, typename RO>
__device__ __forceinline__ void w96d_xchg(cx<double> (&v)[12], double* wr, const double* rd, WO wo, RO ro) {
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 12; ++i) wr[wo(i)] = part ? v[i].y : v[i].x;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      const double x = rd[ro(i)];
      if (part) v[i].y = x;
      else v[i].x = x;
    }
  }
}

template <int SIGN, bool PIN>
__device__ __forceinline__ void w96d_inv_row(cx<double> (&v)[12], double* row, int l,
                                             const cx<double>* __restrict__ tw96, const cx<double>* __restrict__ twg) {
  constexpr int H1 = 52, B2 = 25;
  dft12<double, SIGN, PIN, false>(v);
  const cx<double>* tl = tw96 + 12 * l;
#pragma unroll
  for (int s = 1; s < 12; ++s) v[s] = v[s] * (SIGN > 0 ? tl[s] : conjg(tl[s]));
  const int b = l & 3, hh = l >> 2;
  // step 1 -> 2: B_j[s] at j + 8 (s % 6) + H1 (s / 6); lane (h, b) reads B_b (even i) and
  // B_(b+4) (odd i) of s = 6h + i / 2
  w96d_xchg(v, row + l, row + b + H1 * hh, [](int s) { return 8 * (s % 6) + H1 * (s / 6); },
            [](int i) { return (i & 1) * 4 + 8 * (i >> 1); });
  const cx<double> w8b = SIGN > 0 ? twg[12 * b] : conjg(twg[12 * b]);
#pragma unroll
  for (int sg = 0; sg < 6; ++sg) {
    const cx<double> x0 = v[2 * sg], x1 = v[2 * sg + 1];
    v[2 * sg] = x0 + x1;
    v[2 * sg + 1] = (x0 - x1) * w8b;
  }
  // step 2 -> 3: (h, b, q) at 4h + b + 8 (q % 3) + B2 (q / 3); lane (h, b') reads b = i % 4
  // of the pair q = 3b' + i / 4
  w96d_xchg(v, row + 4 * hh + b, row + 4 * hh + B2 * b, [](int q) { return 8 * (q % 3) + B2 * (q / 3); },
            [](int i) { return (i & 3) + 8 * (i >> 2); });
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    cx<double> q4[4] = {v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]};
    dft<double, 4, SIGN>(q4);
#pragma unroll
    for (int d = 0; d < 4; ++d) v[4 * u + d] = q4[d];
  }
}

template <int SIGN, bool POUT>
__device__ __forceinline__ void w96d_fwd_row(cx<double> (&v)[12], double* row, int l,
                                             const cx<double>* __restrict__ tw96) {
  constexpr int H1 = 52, B2 = 25;
  const int b = l & 3, hh = l >> 2;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    cx<double> q4[4] = {v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]};
    dft<double, 4, SIGN>(q4);
    const bool c1 = ((b + u) & 1) != 0;
    const cx<double> t1 = w8_mul<double, SIGN, 1>(q4[1]), t2 = w8_mul<double, SIGN, 2>(q4[2]),
                     t3 = w8_mul<double, SIGN, 3>(q4[3]);
    v[4 * u] = q4[0];
    v[4 * u + 1] = sel_cx(c1, t1, q4[1]);
    v[4 * u + 2] = sel_cx(c1, t2, q4[2]);
    v[4 * u + 3] = sel_cx(c1, t3, q4[3]);
  }
  // step 3 -> 2 (the inverse's read pattern as writes, and back)
  w96d_xchg(v, row + 4 * hh + B2 * b, row + 4 * hh + b, [](int i) { return (i & 3) + 8 * (i >> 2); },
            [](int q) { return 8 * (q % 3) + B2 * (q / 3); });
#pragma unroll
  for (int sg = 0; sg < 6; ++sg) {
    const cx<double> z0 = v[2 * sg], z1 = v[2 * sg + 1];
    v[2 * sg] = z0 + z1;        // U_b[6h + sg]
    v[2 * sg + 1] = z0 - z1;    // U_(b+4)[6h + sg]
  }
  // step 2 -> 1
  w96d_xchg(v, row + b + H1 * hh, row + l, [](int i) { return (i & 1) * 4 + 8 * (i >> 1); },
            [](int s) { return 8 * (s % 6) + H1 * (s / 6); });
  const cx<double>* tl = tw96 + 12 * l;
#pragma unroll
  for (int s = 1; s < 12; ++s) v[s] = v[s] * (SIGN > 0 ? tl[s] : conjg(tl[s]));
  dft12<double, SIGN, false, POUT>(v);
}

__device__ __forceinline__ void plane_w96d_body(SphereMaps m, const cx<double>* __restrict__ twg,
                                                const double* __restrict__ V, cx<double>* __restrict__ W,
                                                const int bx, const int by) {
  constexpr int N = 96, S = kW96dStride, NB = 48;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int z = bx, b = by;
  const cx<double>* tw96 = twg + N;
  cx<double>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  const int jc = l >> 3, xb = 8 * wave + (l & 7);
  cx<double>* col = tile + xb;
  const int* bm = m.bandmap + jc * NB + xb;
  cx<double> v[12];
  {
    int c[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = bm[8 * NB * i];
    cx<double> e[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) e[i] = wrow[c[i] < 0 ? 0 : c[i]];
#pragma unroll
    for (int r = 0; r < 12; ++r) v[r] = mk<double>(0, 0);
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i < 3 ? i : i + 6] = c[i] < 0 ? mk<double>(0, 0) : e[i];
    w96_inv<double, +1, true, false, S>(v, col, jc, tw96, twg);
    wave_lds_sync();
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int d = 0; d < 4; ++d) col[w96_k(jc, u, d) * S] = v[4 * u + d];
  }
  __syncthreads();
#pragma unroll 1
  for (int rd = 0; rd < 2; ++rd) {
    const int jr = w96d_opaque(l & 7);
    const int y = 8 * wave + (l >> 3) + 48 * rd;
    double vr[12];
    const double* Vrow = V + ((size_t)z * N + y) * N;
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int d = 0; d < 4; ++d) vr[4 * u + d] = Vrow[w96_k(jr, u, d)];
    cx<double>* row = tile + y * S;
    // band column jr + 8r of the row is at position jr + 8r' (r' = r below 3, r - 6 above 8)
#pragma unroll
    for (int r = 0; r < 12; ++r) v[r] = (r < 3 || r >= 9) ? row[jr + 8 * (r < 3 ? r : r - 6)] : mk<double>(0, 0);
    wave_lds_sync();
    w96d_inv_row<+1, true>(v, reinterpret_cast<double*>(row), jr, tw96, twg);
#pragma unroll
    for (int i = 0; i < 12; ++i) v[i] = vr[i] * v[i];
    w96d_fwd_row<-1, true>(v, reinterpret_cast<double*>(row), jr, tw96);
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 12; ++r)
      if (r < 3 || r >= 9) row[jr + 8 * (r < 3 ? r : r - 6)] = v[r];
  }
  __syncthreads();
  {
    const int jf = w96d_opaque(jc);
    const int* bmf = m.bandmap + jf * NB + xb;
    int c[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) c[i] = bmf[8 * NB * i];
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int d = 0; d < 4; ++d) v[4 * u + d] = col[w96_k(jf, u, d) * S];
    w96_fwd<double, -1, true, false, S>(v, col, jf, tw96);
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if (c[i] >= 0) wrow[c[i]] = v[i < 3 ? i : i + 6];
  }
}

__global__ __launch_bounds__(384, 3) void k_plane_w96d(SphereMaps m, const cx<double>* __restrict__ twg,
                                                    const double* __restrict__ V, cx<double>* __restrict__ W) {
  plane_w96d_body(m, twg, V, W, blockIdx.x, blockIdx.y);
}

// FP64 64 x 64 planes.  The FP32 layout in FP64 takes 74 KB per plane (two per CU, the
// kernel LDS-bound); this one keeps only the 32 band columns of each row (36 complex = 72
// eight-byte words per row, 37 KB per plane: four per CU).  The rows exchange the real
// parts and then the imaginary parts through their own words, slot 9s + j (the row
// stride, 72 words, is 8 mod 16, so the two rows of a 16-lane group use different bank
// halves); the columns exchange complex values in place, slot j + 8s of the column's 64
// rows (a bank cycle of 16-byte words serves the 8 consecutive columns of a wave).  A
// workgroup is 4 waves: the 32 band columns, then the 64 rows in two rounds.
constexpr int kW64dStride = 36;

template <int SIGN, bool PIN, bool POUT>
__device__ __forceinline__ void w64d_col(cx<double> (&v)[8], cx<double>* col, int j, const cx<double> (&tw)[8]) {
  constexpr int S = kW64dStride;
  if constexpr (PIN) dft8_in0167<double, SIGN>(v);
  else dft<double, 8, SIGN>(v);
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = v[r] * (SIGN > 0 ? tw[r] : conjg(tw[r]));
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) col[(j + 8 * r) * S] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = col[(8 * j + i) * S];
  if constexpr (POUT) dft8_out0167<double, SIGN>(v);
  else dft<double, 8, SIGN>(v);
}

template <int SIGN, bool PIN, bool POUT>
__device__ __forceinline__ void w64d_row(cx<double> (&v)[8], double* row, int j, const cx<double> (&tw)[8]) {
  if constexpr (PIN) dft8_in0167<double, SIGN>(v);
  else dft<double, 8, SIGN>(v);
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = v[r] * (SIGN > 0 ? tw[r] : conjg(tw[r]));
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) row[9 * r + j] = part ? v[r].y : v[r].x;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double x = row[9 * j + i];
      if (part) v[i].y = x;
      else v[i].x = x;
    }
  }
  if constexpr (POUT) dft8_out0167<double, SIGN>(v);
  else dft<double, 8, SIGN>(v);
}

__device__ __forceinline__ void plane_w64d_body(SphereMaps m, const cx<double>* __restrict__ twg,
                                                const double* __restrict__ V, cx<double>* __restrict__ W,
                                                const int bx, const int by) {
  constexpr int N = 64, S = kW64dStride, NB = 32;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int z = bx, b = by;
  const cx<double>* tw8 = twg + N;
  cx<double>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
  const int jc = l >> 3, xb = 8 * wave + (l & 7);
  cx<double>* col = tile + xb;
  cx<double> v[8], tw[8];
  auto load_tw = [&](int j) {
    tw[0] = mk<double>(1, 0);
#pragma unroll
    for (int r = 1; r < 8; ++r) tw[r] = tw8[8 * j + r];
  };
  // inverse along i2 on the 32 band columns (inputs in the band rows jc + 8q: band index
  // jc + 8q, q = 0..3, at rows jc, jc + 8, jc + 48, jc + 56)
  {
    const int* bm = m.bandmap + jc * NB + xb;
    int c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = bm[8 * NB * q];
    load_tw(jc);
    cx<double> e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) e[q] = wrow[c[q] < 0 ? 0 : c[q]];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q < 2 ? q : q + 4] = c[q] < 0 ? mk<double>(0, 0) : e[q];
    v[2] = v[3] = v[4] = v[5] = mk<double>(0, 0);
    w64d_col<+1, true, false>(v, col, jc, tw);
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 8; ++t) col[(jc + 8 * t) * S] = v[t];
  }
  __syncthreads();
  // rows in two rounds: inverse along i1, times V, forward along i1 keeping the band
  // columns (band column jr + 8r of the row is at position jr + 8r', r' = r below 2,
  // r - 4 above 5)
#pragma unroll 1
  for (int rd = 0; rd < 2; ++rd) {
    const int jr = w96d_opaque(l & 7);
    const int y = 8 * wave + (l >> 3) + 32 * rd;
    double vr[8];
    const double* Vrow = V + ((size_t)z * N + y) * N;
#pragma unroll
    for (int t = 0; t < 8; ++t) vr[t] = Vrow[jr + 8 * t];
    load_tw(jr);
    cx<double>* row = tile + y * S;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = (r < 2 || r >= 6) ? row[jr + 8 * (r < 2 ? r : r - 4)] : mk<double>(0, 0);
    wave_lds_sync();
    w64d_row<+1, true, false>(v, reinterpret_cast<double*>(row), jr, tw);
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = vr[t] * v[t];
    w64d_row<-1, false, true>(v, reinterpret_cast<double*>(row), jr, tw);
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (t < 2 || t >= 6) row[jr + 8 * (t < 2 ? t : t - 4)] = v[t];
  }
  __syncthreads();
  // forward along i2 on the band columns, keeping the band rows, written to W
  {
    const int jf = w96d_opaque(jc);
    const int* bm = m.bandmap + jf * NB + xb;
    int c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = bm[8 * NB * q];
    load_tw(jf);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = col[(jf + 8 * r) * S];
    w64d_col<-1, false, true>(v, col, jf, tw);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c[q] >= 0) wrow[c[q]] = v[q < 2 ? q : q + 4];
  }
}

__global__ __launch_bounds__(256, 4) void k_plane_w64d(SphereMaps m, const cx<double>* __restrict__ twg,
                                                       const double* __restrict__ V, cx<double>* __restrict__ W) {
  plane_w64d_body(m, twg, V, W, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Density pass (EvalRhoPsi.cpp:138-150): per (i3, band group) plane, inverse FFT along i2
// on the active columns and along i1, then sum w_b |psi_b(r)|^2 over the group's bands
// ---------------------------------------------------------------------------
constexpr int kRhoThreads = 256;

// part[g][i3][i2][i1] = sum over bands b = g, g + groups, ... of wts[b] |psi_b|^2 (FP64)
template <int N1, int N2>
__global__ __launch_bounds__(kRhoThreads) void k_plane_rho(SphereMaps m, const cx<double>* __restrict__ tw1g,
                                                           const cx<double>* __restrict__ tw2g,
                                                           const cx<double>* __restrict__ W, int nb, int groups,
                                                           const double* __restrict__ wts,
                                                           double* __restrict__ part) {
  constexpr int NT = kRhoThreads;
  constexpr int S1 = plane_stride<N1>();
  constexpr int PTS = (N1 * N2 + NT - 1) / NT;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  cx<double>* tw1 = tile + N2 * S1;
  cx<double>* tw2 = tw1 + N1;
  int* xa = reinterpret_cast<int*>(tw2 + N2);
  unsigned char* xact = reinterpret_cast<unsigned char*>(xa + N1);
  const int tid = threadIdx.x;
  const int z = blockIdx.x, g0 = blockIdx.y;
  const int nxa = m.nxa;
  for (int i = tid; i < N1; i += NT) { tw1[i] = tw1g[i]; xact[i] = 0; }
  for (int i = tid; i < N2; i += NT) tw2[i] = tw2g[i];
  __syncthreads();
  for (int i = tid; i < nxa; i += NT) { const int x = m.xa[i]; xa[i] = x; xact[x] = 1; }
  double acc[PTS];
#pragma unroll
  for (int q = 0; q < PTS; ++q) acc[q] = 0.0;
  const XLiveHooks hk{xact};
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
    lds_fft<double, N2, +1, NT, N1, true>(tile, nxa, S1, [&](int t) { return xa[t]; }, tw2, tid);
    lds_fft<double, N1, +1, NT, N2, true>(tile, N2, 1, [](int t) { return t * S1; }, tw1, tid, hk);
    const double wb = wts[b];
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      const int idx = tid + q * NT;
      if (idx < N1 * N2) {
        const int y = idx / N1, x = idx - y * N1;
        const cx<double> v = tile[y * S1 + x];
        acc[q] += wb * (v.x * v.x + v.y * v.y);
      }
    }
  }
  double* out = part + ((size_t)g0 * m.nr3 + z) * (size_t)(N1 * N2);
#pragma unroll
  for (int q = 0; q < PTS; ++q) {
    const int idx = tid + q * NT;
    if (idx < N1 * N2) out[idx] = acc[q];
  }
}

// The same with band pruning (the sphere's i1, i2 in [-n/4, n/4)): inverse along i2 on
// the n/2 band columns and along i1 on every row, both with the first stage's out-of-band
// inputs known zero; w_b |psi_b|^2 is added in registers at the last row stage.
template <int N, int NT>
__global__ __launch_bounds__(NT) void k_plane_rho_band(SphereMaps m, const cx<double>* __restrict__ twg,
                                                       const cx<double>* __restrict__ W, int nb, int groups,
                                                       const double* __restrict__ wts,
                                                       double* __restrict__ part) {
  constexpr int S1 = N + 1, Q = N / 4, NBC = N / 2;
  static_assert(N % 4 == 0, "band pruning needs n % 4 == 0");
  extern __shared__ __align__(16) unsigned char smem[];
  cx<double>* tile = reinterpret_cast<cx<double>*>(smem);
  cx<double>* tw = tile + N * S1;
  const int tid = threadIdx.x;
  const int z = blockIdx.x, g0 = blockIdx.y;
  // outputs of the last inverse row stage: lane slot q, element r -> row t, column j + r NSL
  constexpr int RL = radix_at(N, nstages(N) - 1, false), NSL = N / RL;
  constexpr int TOTL = N * NSL, MBL = (TOTL + NT - 1) / NT;
  double acc[MBL][RL];
#pragma unroll
  for (int q = 0; q < MBL; ++q)
#pragma unroll
    for (int r = 0; r < RL; ++r) acc[q][r] = 0.0;
  auto band_x = [](int t) { return t < Q ? t : t + NBC; };
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  auto col_base = [&](int t) { return band_x(t); };
  auto row_base = [](int t) { return t * S1; };
  const NoPost none;
  for (int b = g0; b < nb; b += groups) {
    __syncthreads();
    for (int i = tid; i < NBC * NBC; i += NT) {
      const int yy = i / NBC, xx = i - yy * NBC;
      tile[band_x(yy) * S1 + band_x(xx)] = mk<double>(0, 0);
    }
    __syncthreads();
    const cx<double>* wrow = W + ((size_t)b * m.nr3 + z) * m.ncol_pad;
    for (int c = tid; c < m.ncol; c += NT) {
      const int xy = m.colxy[c];
      tile[(xy / N) * S1 + xy % N] = wrow[c];
    }
    __syncthreads();
    const double wb = wts[b];
    auto accum = [&](int q, int r, cx<double>& v) { acc[q][r] += wb * (v.x * v.x + v.y * v.y); };
    band_fft<double, N, +1, false, NT, NBC, S1, true, false>(tile, col_base, tw, twg, tid, none);
    band_fft<double, N, +1, false, NT, N, 1, true, false>(tile, row_base, tw, twg, tid, accum);
  }
  double* out = part + ((size_t)g0 * m.nr3 + z) * (size_t)(N * N);
#pragma unroll
  for (int q = 0; q < MBL; ++q) {
    const int idx = tid + q * NT;
    if (TOTL % NT == 0 || idx < TOTL) {
      const int j = idx / N, t = idx - j * N;
#pragma unroll
      for (int r = 0; r < RL; ++r) out[t * N + j + r * NSL] = acc[q][r];
    }
  }
}

// rho[i] = sum_g part[g][i], groups summed in order (deterministic)
__global__ void k_sum_groups(const double* __restrict__ part, int groups, size_t n, double* __restrict__ rho) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int g = 0; g < groups; ++g) s += part[(size_t)g * n + i];
  rho[i] = s;
}

// ---------------------------------------------------------------------------
// Pass 3: forward transform along i3, gather the sphere, fused epilogue
// ---------------------------------------------------------------------------
template <typename T, int NZ, int CT, int NT, int PE>
__device__ __forceinline__ void zfwd_body(SphereMaps m, const cx<T>* __restrict__ tw3g,
                                             const cx<T>* __restrict__ W,
                                             const T* __restrict__ kin,
                                             const cx<T>* __restrict__ X,
                                             const cx<T>* __restrict__ Hnl,
                                             const cx<T>* __restrict__ Z,
                                             cx<T>* __restrict__ out, int ld, T inv_nnr,
                                             T alpha, T beta, T gamma, const int bx, const int by) {
  __shared__ cx<T> tile[NZ * CT];
  __shared__ cx<T> tw[NZ];
  constexpr int TWP = (NZ + NT - 1) / NT;
  constexpr int WP = (NZ * CT + NT - 1) / NT;
  const int tid = threadIdx.x;
  const int c0 = bx * CT;
  const int b = by;
  const int cn = min(CT, m.ncol_pad - c0);
  const int kb = m.colstart[c0], ke = m.colstart[c0 + cn];
  const size_t ob = (size_t)b * ld;
  // Loads use clamped, always valid indices and no branch, so they issue back to back;
  // out-of-range lanes are discarded afterwards.  Register copies are separate real
  // arrays (arrays of complex structs end up in scratch).
  T twr[TWP], twi[TWP];
#pragma unroll
  for (int j = 0; j < TWP; ++j) {
    const cx<T> v = tw3g[min(tid + j * NT, NZ - 1)];
    twr[j] = v.x; twi[j] = v.y;
  }
  const cx<T>* wb = W + (size_t)b * NZ * m.ncol_pad + c0;
  T wr[WP], wi[WP];
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    const int i = min(tid + j * NT, NZ * CT - 1);
    const int z = i / CT, c = min(i - z * CT, cn - 1);
    const cx<T> v = wb[(size_t)z * m.ncol_pad + c];
    wr[j] = v.x; wi[j] = v.y;
  }
  // epilogue operands, loaded before the transform and combined after it:
  // out = alpha/nnr * FFT + (alpha * kin + beta) * X + alpha * Hnl + gamma * Z; a null
  // Hnl or Z reads X again with a zero coefficient
  const cx<T>* hp = Hnl ? Hnl : X;
  const cx<T>* zp = Z ? Z : X;
  const T hc = Hnl ? alpha : T(0), zc = Z ? gamma : T(0);
  const int klast = max(ke - 1, 0);
  T xr[PE], xi[PE], hr[PE], hi[PE], zr[PE], zi[PE], kv[PE];
  int kzv[PE], kcv[PE];
#pragma unroll
  for (int j = 0; j < PE; ++j) {
    const int k = min(kb + tid + j * NT, klast);
    const cx<T> x = X[ob + k], h = hp[ob + k], zz = zp[ob + k];
    xr[j] = x.x; xi[j] = x.y;
    hr[j] = h.x; hi[j] = h.y;
    zr[j] = zz.x; zi[j] = zz.y;
    kv[j] = kin[k];
    kzv[j] = m.kz[k];
    kcv[j] = m.kcol[k];
  }
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    const int i = tid + j * NT;
    if (i < NZ * CT) tile[i] = ((i % CT) < cn) ? mk<T>(wr[j], wi[j]) : mk<T>(0, 0);
  }
#pragma unroll
  for (int j = 0; j < TWP; ++j)
    if (tid + j * NT < NZ) tw[tid + j * NT] = mk<T>(twr[j], twi[j]);
  __syncthreads();
  lds_fft<T, NZ, -1, NT, CT, true>(tile, CT, CT, [](int t) { return t; }, tw, tid);
  const T an = alpha * inv_nnr;
#pragma unroll
  for (int j = 0; j < PE; ++j) {
    const int k = kb + tid + j * NT;
    if (k < ke) {
      const T cxk = alpha * kv[j] + beta;
      const cx<T> t = tile[kzv[j] * CT + (kcv[j] - c0)];
      cx<T> o = an * t + cxk * mk<T>(xr[j], xi[j]);
      o = o + hc * mk<T>(hr[j], hi[j]);
      o = o + zc * mk<T>(zr[j], zi[j]);
      out[ob + k] = o;
    }
  }
  for (int k = kb + PE * NT + tid; k < ke; k += NT) {
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

template <typename T, int NZ, int CT, int NT, int PE>
__global__ __launch_bounds__(NT) void k_zfwd(SphereMaps m, const cx<T>* __restrict__ tw3g,
                                             const cx<T>* __restrict__ W,
                                             const T* __restrict__ kin,
                                             const cx<T>* __restrict__ X,
                                             const cx<T>* __restrict__ Hnl,
                                             const cx<T>* __restrict__ Z,
                                             cx<T>* __restrict__ out, int ld, T inv_nnr,
                                             T alpha, T beta, T gamma) {
  zfwd_body<T, NZ, CT, NT, PE>(m, tw3g, W, kin, X, Hnl, Z, out, ld, inv_nnr, alpha, beta, gamma, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Dense 3D transform (fft_fftw3.cpp:13-46): one pass per axis over every line
// ---------------------------------------------------------------------------
// The array is [outer][N][inner]; a line is the N elements at stride `inner`.  A block
// transforms kColTile lines: consecutive lines when inner > 1 (coalesced along inner),
// or consecutive contiguous lines when inner == 1.  `scale` multiplies the result.
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(kColThreads) void k_fft_lines(cx<T>* __restrict__ data, int inner, int nlines,
                                                           const cx<T>* __restrict__ twg, T scale) {
  constexpr int CT = kColTile, NT = kColThreads, S = CT + 1;
  __shared__ cx<T> tile[N * S];
  __shared__ cx<T> tw[N];
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
    tile[e * S + t] = (l0 + t < nlines) ? data[addr(t, e)] : mk<T>(0, 0);
  }
  __syncthreads();
  lds_fft<T, N, SIGN, NT, CT, true>(tile, CT, S, [](int t) { return t; }, tw, tid);
  for (int idx = tid; idx < N * CT; idx += NT) {
    int t, e;
    if (inner == 1) { t = idx / N; e = idx - t * N; }
    else { e = idx / CT; t = idx - e * CT; }
    if (l0 + t < nlines) data[addr(t, e)] = scale * tile[e * S + t];
  }
}

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
#define PWH_FFT_SIZES(X) \
  X(12) X(16) X(18) X(20) X(24) X(30) X(32) X(36) X(40) X(48) X(50) X(54) X(60) X(64) \
  X(72) X(80) X(90) X(96) X(100) X(108) X(120) X(128)

bool fft_size_supported(int n) {
  switch (n) {
#define PWH_CASE(N) case N: return true;
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return false;
  }
}

bool plane_supported(int n1, int n2) { return n1 == n2 && fft_size_supported(n1); }

// the compiled plane kernel of this precision fits the LDS of one CU
static bool plane_compiled_fits(const SphereMaps& m, bool fp64) {
  if (!plane_supported(m.nr1, m.nr2)) return false;
  const int n = m.nr1;
  const size_t cxb = fp64 ? 16 : 8;
  const size_t lds = m.band_xy && n % 4 == 0 ? cxb * ((size_t)n * (n + 1) + n)
                                               : cxb * ((size_t)n * (n + 1) + 2 * n) + 5 * (size_t)n;
  return lds <= 160 * 1024;
}

bool plane_uses_split(const SphereMaps& m, bool fp64) {
  if (m.path == FFT_COMPILED && plane_compiled_fits(m, fp64)) return false;
  return m.path == FFT_SPLIT || !plane_rt_fits(m.nr1, m.nr2, fp64);
}

bool compiled_plans_cover(int n1, int n2, int n3) { return plane_supported(n1, n2) && fft_size_supported(n3); }

template <typename T, int N, int NT>
static hipError_t launch_plane_lds(const SphereMaps& m, const void* tw1, const void* tw2, const T* V,
                                   cx<T>* W, int nb, hipStream_t s) {
  constexpr size_t smem = plane_smem<T, N, N>();
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane<T, N, N, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  dim3 grid(m.nr3, nb);
  hipLaunchKernelGGL((k_plane<T, N, N, NT>), grid, dim3(NT), smem, s, m,
                     static_cast<const cx<T>*>(tw1), static_cast<const cx<T>*>(tw2), V, W);
  return hipGetLastError();
}

// Plane pass dispatch: the band-pruned kernel when the sphere's columns lie in the band,
// else the LDS Stockham kernel with kPlaneThreads (512) lanes (measured faster than 256
// at Si64).
static bool w64_enabled() {
  static const bool on = [] {
    const char* e = getenv("PWH_PLANE_W64");
    return !(e && e[0] == '0');
  }();
  return on;
}
static bool w64d_enabled() {
  static const bool on = [] {
    const char* e = getenv("PWH_PLANE_W64D");
    return !(e && e[0] == '0');
  }();
  return on;
}
static bool w96_enabled() {
  static const bool on = [] {
    const char* e = getenv("PWH_PLANE_W96");
    return !(e && e[0] == '0');
  }();
  return on;
}
static bool band_enabled() {
  static const bool on = [] {
    const char* e = getenv("PWH_PLANE_BAND");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Dynamic LDS of one plane workgroup when at most `cap` of them may share a CU (cap 0: no
// cap, the kernel's own LDS decides).  The padding leaves LDS and wave slots to the
// nonlocal GEMM workgroups that run beside the plane pass on the second stream: the plane
// pass is VALU-bound and the GEMMs MFMA-bound, so a CU holding both issues from both.  A cap
// the kernel's own LDS already meets leaves it unchanged.
static size_t plane_lds_capped(size_t base, int cap) {
  constexpr size_t lds = 160 * 1024;
  if (cap <= 0) return base;
  const size_t want = (lds / (size_t)(cap + 1) / 1024 + 1) * 1024;
  const size_t sm = base > want ? base : want;
  return (size_t)cap * sm <= lds ? sm : base;
}

// one-time raise of a kernel's dynamic LDS limit to the whole CU (the capped sizes); one
// instance, and so one flag, per kernel
template <auto KERN>
static hipError_t plane_lds_attr() {
  static const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(KERN),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  return e;
}

template <typename T, int N>
static hipError_t launch_plane_n(const SphereMaps& m, const void* tw1, const void* tw2, const T* V,
                                 cx<T>* W, int nb, hipStream_t s, int cap) {
  const bool w64_on = w64_enabled();
  if constexpr (N == 64 && sizeof(T) == 8) {
    if (m.band_xy && w64_on && w64d_enabled()) {
      const size_t sm = plane_lds_capped(sizeof(cx<double>) * (size_t)N * kW64dStride, cap);
      if (hipError_t e = plane_lds_attr<&k_plane_w64d>(); e != hipSuccess) return e;
      hipLaunchKernelGGL(k_plane_w64d, dim3(m.nr3, nb), dim3(256), sm, s, m,
                         reinterpret_cast<const cx<double>*>(tw1), reinterpret_cast<const double*>(V),
                         reinterpret_cast<cx<double>*>(W));
      return hipGetLastError();
    }
  }
  if constexpr (N == 64) {
    if (m.band_xy && w64_on) {
      const size_t sm = plane_lds_capped(sizeof(cx<T>) * (size_t)N * kW64Stride, cap);
      if (hipError_t e = plane_lds_attr<&k_plane_w64<T>>(); e != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_w64<T>), dim3(m.nr3, nb), dim3(512), sm, s, m, static_cast<const cx<T>*>(tw1),
                         V, W);
      return hipGetLastError();
    }
  }
  if constexpr (N == 96 && sizeof(T) == 4) {
    if (m.band_xy && w96_enabled()) {
      const size_t sm = plane_lds_capped(sizeof(cx<T>) * (size_t)N * kW96Stride, cap);
      if (hipError_t e = plane_lds_attr<&k_plane_w96<T>>(); e != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_w96<T>), dim3(m.nr3, nb), dim3(768), sm, s, m, static_cast<const cx<T>*>(tw1),
                         V, W);
      return hipGetLastError();
    }
  }
  if constexpr (N == 96 && sizeof(T) == 8) {
    if (m.band_xy && w96_enabled()) {
      const size_t sm = plane_lds_capped(sizeof(cx<double>) * (size_t)N * kW96dStride, cap);
      if (hipError_t e = plane_lds_attr<&k_plane_w96d>(); e != hipSuccess) return e;
      hipLaunchKernelGGL(k_plane_w96d, dim3(m.nr3, nb), dim3(384), sm, s, m,
                         reinterpret_cast<const cx<double>*>(tw1), reinterpret_cast<const double*>(V),
                         reinterpret_cast<cx<double>*>(W));
      return hipGetLastError();
    }
  }
  const bool band_on = band_enabled();
  if constexpr (N % 4 == 0 && band_plane_smem<T, N>() <= 160 * 1024) {
    if (m.band_xy && band_on) {
      constexpr int NT = band_plane_threads<N>();
      const size_t bsm = plane_lds_capped(band_plane_smem<T, N>(), cap);
      if (hipError_t e = plane_lds_attr<&k_plane_band<T, N, NT>>(); e != hipSuccess) return e;
      dim3 grid(m.nr3, nb);
      hipLaunchKernelGGL((k_plane_band<T, N, NT>), grid, dim3(NT), bsm, s, m,
                         static_cast<const cx<T>*>(tw1), V, W);
      return hipGetLastError();
    }
  }
  constexpr size_t smem = plane_smem<T, N, N>();
  if constexpr (smem > 160 * 1024) {
    return hipErrorInvalidValue;
  } else {
    return launch_plane_lds<T, N, kPlaneThreads>(m, tw1, tw2, V, W, nb, s);
  }
}

static int col_tile_index(int ct) { return ct == 8 ? 0 : ct == 16 ? 1 : ct == 32 ? 2 : 3; }

// prefetch depth: entries per lane that cover the largest tile, capped at 4
static int col_pe(const SphereMaps& m, int ct, int nt) {
  const int per = (m.tile_max[col_tile_index(ct)] + nt - 1) / nt;
  return per <= 1 ? 1 : per <= 2 ? 2 : 4;
}

template <typename T, int N, int CT, int NT, int PE>
static hipError_t zinv_pe(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W, int nb,
                          hipStream_t s) {
  dim3 grid((m.ncol_pad + CT - 1) / CT, nb);
  hipLaunchKernelGGL((k_zinv<T, N, CT, NT, PE>), grid, dim3(NT), 0, s, m, static_cast<const cx<T>*>(tw3), X,
                     ldx, W);
  return hipGetLastError();
}

template <typename T, int N, int CT, int NT>
static hipError_t zinv_shape(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W, int nb,
                             hipStream_t s) {
  switch (col_pe(m, CT, NT)) {
    case 1: return zinv_pe<T, N, CT, NT, 1>(m, tw3, X, ldx, W, nb, s);
    case 2: return zinv_pe<T, N, CT, NT, 2>(m, tw3, X, ldx, W, nb, s);
    default: return zinv_pe<T, N, CT, NT, 4>(m, tw3, X, ldx, W, nb, s);
  }
}

// Launch shapes measured at the 64^3 Si64 and 96^3 Si216 blocks (FP32 and FP64): the
// gather-and-write pass runs best as one-wave blocks of 8 columns up to 64-point columns
// and as two-wave blocks of 16 columns above (96: FP32 0.187 -> 0.119 ms, FP64 0.243 ->
// 0.219 ms; the FP32 rows of 16 columns are whole 128-byte lines of W); the
// transform-and-epilogue pass as 256-lane blocks of 32 columns (16 x 128 when that tile
// would exceed 32 KB of LDS).  The column transforms are the same arithmetic in either
// shape, so the result does not depend on it.
template <typename T, int N>
static hipError_t zinv_n(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W, int nb,
                         hipStream_t s) {
  if constexpr (N > 64) return zinv_shape<T, N, 16, 128>(m, tw3, X, ldx, W, nb, s);
  else return zinv_shape<T, N, 8, 64>(m, tw3, X, ldx, W, nb, s);
}

template <typename T>
hipError_t launch_zinv(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W,
                       int nb, hipStream_t s) {
  if (m.path != FFT_COMPILED || !fft_size_supported(m.nr3)) return launch_zinv_rt<T>(m, tw3, X, ldx, W, nb, s);
  switch (m.nr3) {
#define PWH_CASE(N) \
  case N: return zinv_n<T, N>(m, tw3, X, ldx, W, nb, s);
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_plane(const SphereMaps& m, const void* tw1, const void* tw2, const T* V, cx<T>* W,
                        int nb, cx<T>* Y, int chunk, hipStream_t s, int cap) {
  if (m.path != FFT_COMPILED || !plane_compiled_fits(m, sizeof(T) == sizeof(double)))
    return launch_plane_rt<T>(m, tw1, tw2, V, W, nb, Y, chunk, s);
  switch (m.nr1) {
#define PWH_CASE(N) \
  case N: return launch_plane_n<T, N>(m, tw1, tw2, V, W, nb, s, cap);
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
}

struct ZfwdArgs {
  const void* tw3;
  const void* W;
  const void* kin;
  const void* X;
  const void* Hnl;
  const void* Z;
  void* out;
  int ld, nb;
  double inv_nnr, alpha, beta, gamma;
};

template <typename T, int N, int CT, int NT, int PE>
static hipError_t zfwd_pe(const SphereMaps& m, const ZfwdArgs& a, hipStream_t s) {
  dim3 grid((m.ncol_pad + CT - 1) / CT, a.nb);
  hipLaunchKernelGGL((k_zfwd<T, N, CT, NT, PE>), grid, dim3(NT), 0, s, m, static_cast<const cx<T>*>(a.tw3),
                     static_cast<const cx<T>*>(a.W), static_cast<const T*>(a.kin),
                     static_cast<const cx<T>*>(a.X), static_cast<const cx<T>*>(a.Hnl),
                     static_cast<const cx<T>*>(a.Z), static_cast<cx<T>*>(a.out), a.ld, (T)a.inv_nnr, (T)a.alpha,
                     (T)a.beta, (T)a.gamma);
  return hipGetLastError();
}

template <typename T, int N, int CT, int NT>
static hipError_t zfwd_shape(const SphereMaps& m, const ZfwdArgs& a, hipStream_t s) {
  switch (col_pe(m, CT, NT)) {
    case 1: return zfwd_pe<T, N, CT, NT, 1>(m, a, s);
    case 2: return zfwd_pe<T, N, CT, NT, 2>(m, a, s);
    default: return zfwd_pe<T, N, CT, NT, 4>(m, a, s);
  }
}

template <typename T, int N>
static hipError_t zfwd_n(const SphereMaps& m, const ZfwdArgs& a, hipStream_t s) {
  if constexpr ((size_t)N * 32 * sizeof(cx<T>) <= 32 * 1024) return zfwd_shape<T, N, 32, 256>(m, a, s);
  else return zfwd_shape<T, N, 16, 128>(m, a, s);
}

template <typename T>
hipError_t launch_zfwd_epilogue(const SphereMaps& m, const void* tw3, const cx<T>* W, const T* kin,
                                const cx<T>* X, const cx<T>* Hnl, const cx<T>* Z, cx<T>* out,
                                int ld, int nb, T inv_nnr, T alpha, T beta, T gamma,
                                hipStream_t s) {
  if (m.path != FFT_COMPILED || !fft_size_supported(m.nr3))
    return launch_zfwd_rt<T>(m, tw3, W, kin, X, Hnl, Z, out, ld, nb, inv_nnr, alpha, beta, gamma, s);
  const ZfwdArgs a{tw3, W, kin, X, Hnl, Z, out, ld, nb, (double)inv_nnr, (double)alpha, (double)beta,
                   (double)gamma};
  switch (m.nr3) {
#define PWH_CASE(N) \
  case N: return zfwd_n<T, N>(m, a, s);
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Grouped passes over a batch.  The systems whose pass runs the same compiled kernel (the
// same length and plane variant) share one launch; block L belongs to the system whose
// first block blk0 is the largest one <= L and runs the one-system body with its local
// block.  The column passes prefetch PE = 4 entries per lane for every system: PE only
// decides which entries are loaded before the first barrier, so results do not change.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ const GroupHpsi<T>& group_sys(const GroupHpsi<T>* __restrict__ g, int n, int L) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].blk0 <= L) lo = mid;
    else hi = mid - 1;
  }
  return g[lo];
}

template <typename T, int NZ, int CT, int NT>
__global__ __launch_bounds__(NT) void k_zinv_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0, ncb = (p.m.ncol_pad + CT - 1) / CT;
  zinv_body<T, NZ, CT, NT, 4>(p.m, p.tw3, p.X, p.ld, p.W, local % ncb, local / ncb);
}

template <typename T, int NZ, int CT, int NT>
__global__ __launch_bounds__(NT) void k_zfwd_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0, ncb = (p.m.ncol_pad + CT - 1) / CT;
  zfwd_body<T, NZ, CT, NT, 4>(p.m, p.tw3, p.W, p.kin, p.X, p.Hnl, p.Z, p.out, p.ld, p.inv_nnr, p.alpha, p.beta,
                              p.gamma, local % ncb, local / ncb);
}

template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void k_plane_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_body<T, N, N, NT>(p.m, p.tw1, p.tw2, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void k_plane_band_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_band_body<T, N, NT>(p.m, p.tw1, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

template <typename T>
__global__ __launch_bounds__(512) void k_plane_w64_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_w64_body<T>(p.m, p.tw1, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

template <typename T>
__global__ __launch_bounds__(768) void k_plane_w96_gc(const GroupHpsi<T>* __restrict__ g, int n) {
  const GroupHpsi<T>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_w96_body<T>(p.m, p.tw1, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

__global__ __launch_bounds__(256, 4) void k_plane_w64d_gc(const GroupHpsi<double>* __restrict__ g, int n) {
  const GroupHpsi<double>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_w64d_body(p.m, p.tw1, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

__global__ __launch_bounds__(384, 3) void k_plane_w96d_gc(const GroupHpsi<double>* __restrict__ g, int n) {
  const GroupHpsi<double>& p = group_sys(g, n, blockIdx.x);
  const int local = blockIdx.x - p.blk0;
  plane_w96d_body(p.m, p.tw1, p.V, p.W, local % p.m.nr3, local / p.m.nr3);
}

enum GroupKind { GK_RT = 0, GK_COL = 1, GK_W64 = 2, GK_BAND = 3, GK_LDS = 4, GK_W96 = 5 };
constexpr int kGroupKeyN = 4096;   // key = kind * kGroupKeyN + length

template <typename T, int N>
static int plane_group_kind(const SphereMaps& m, size_t* smem) {
  if constexpr (N == 64) {
    if (m.band_xy && w64_enabled()) {
      *smem = sizeof(cx<T>) * (size_t)N * (sizeof(T) == 8 && w64d_enabled() ? kW64dStride : kW64Stride);
      return GK_W64;
    }
  }
  if constexpr (N == 96) {
    if (m.band_xy && w96_enabled()) {
      *smem = sizeof(cx<T>) * (size_t)N * (sizeof(T) == 4 ? kW96Stride : kW96dStride);
      return GK_W96;
    }
  }
  if constexpr (N % 4 == 0 && band_plane_smem<T, N>() <= 160 * 1024) {
    if (m.band_xy && band_enabled()) {
      *smem = band_plane_smem<T, N>();
      return GK_BAND;
    }
  }
  if constexpr (plane_smem<T, N, N>() > 160 * 1024) {
    return -1;
  } else {
    *smem = plane_smem<T, N, N>();
    return GK_LDS;
  }
}

template <typename T>
int group_fft_key(int pass, const SphereMaps& m, int nb, int* blocks, size_t* smem, int* ct) {
  const bool fp64 = std::is_same<T, double>::value;
  *smem = 0;
  *ct = 0;
  if (pass == 0 || pass == 2) {
    if (m.path == FFT_COMPILED && fft_size_supported(m.nr3)) {
      const int c = pass == 0 ? 8 : ((size_t)m.nr3 * 32 * sizeof(cx<T>) <= 32 * 1024 ? 32 : 16);
      *blocks = (m.ncol_pad + c - 1) / c * nb;
      return GK_COL * kGroupKeyN + m.nr3;
    }
  } else if (pass == 1) {
    if (m.path == FFT_COMPILED && plane_compiled_fits(m, fp64)) {
      int kind = -1;
      switch (m.nr1) {
#define PWH_CASE(N) \
  case N: kind = plane_group_kind<T, N>(m, smem); break;
        PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
        default: break;
      }
      if (kind < 0) return -1;
      *blocks = m.nr3 * nb;
      return kind * kGroupKeyN + m.nr1;
    }
  } else {
    return -1;
  }
  int zb, pb;
  size_t zs, ps;
  if (!group_fft_shape<T>(m, nb, ct, &zb, &pb, &zs, &ps)) return -1;
  *blocks = pass == 1 ? pb : zb;
  *smem = pass == 1 ? ps : zs;
  return GK_RT * kGroupKeyN;
}

template <typename K>
static hipError_t group_attr(K kern, size_t smem) {
  if (smem <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)smem);
}

template <typename T, int N>
static hipError_t launch_group_n(int pass, int kind, const GroupHpsi<T>* g, int n, int blocks, hipStream_t s) {
  if (pass == 0) {
    hipLaunchKernelGGL((k_zinv_gc<T, N, 8, 64>), dim3(blocks), dim3(64), 0, s, g, n);
    return hipGetLastError();
  }
  if (pass == 2) {
    if constexpr ((size_t)N * 32 * sizeof(cx<T>) <= 32 * 1024)
      hipLaunchKernelGGL((k_zfwd_gc<T, N, 32, 256>), dim3(blocks), dim3(256), 0, s, g, n);
    else
      hipLaunchKernelGGL((k_zfwd_gc<T, N, 16, 128>), dim3(blocks), dim3(128), 0, s, g, n);
    return hipGetLastError();
  }
  hipError_t e = hipSuccess;
  if constexpr (N == 64 && sizeof(T) == 8) {
    if (kind == GK_W64 && w64d_enabled()) {
      constexpr size_t sm = sizeof(cx<double>) * (size_t)N * kW64dStride;
      hipLaunchKernelGGL(k_plane_w64d_gc, dim3(blocks), dim3(256), sm, s,
                         reinterpret_cast<const GroupHpsi<double>*>(g), n);
      return hipGetLastError();
    }
  }
  if constexpr (N == 64) {
    if (kind == GK_W64) {
      constexpr size_t sm = sizeof(cx<T>) * (size_t)N * kW64Stride;
      if ((e = group_attr(&k_plane_w64_gc<T>, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_w64_gc<T>), dim3(blocks), dim3(512), sm, s, g, n);
      return hipGetLastError();
    }
  }
  if constexpr (N == 96 && sizeof(T) == 4) {
    if (kind == GK_W96) {
      constexpr size_t sm = sizeof(cx<T>) * (size_t)N * kW96Stride;
      if ((e = group_attr(&k_plane_w96_gc<T>, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_w96_gc<T>), dim3(blocks), dim3(768), sm, s, g, n);
      return hipGetLastError();
    }
  }
  if constexpr (N == 96 && sizeof(T) == 8) {
    if (kind == GK_W96) {
      constexpr size_t sm = sizeof(cx<double>) * (size_t)N * kW96dStride;
      if ((e = group_attr(&k_plane_w96d_gc, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL(k_plane_w96d_gc, dim3(blocks), dim3(384), sm, s,
                         reinterpret_cast<const GroupHpsi<double>*>(g), n);
      return hipGetLastError();
    }
  }
  if constexpr (N % 4 == 0 && band_plane_smem<T, N>() <= 160 * 1024) {
    if (kind == GK_BAND) {
      constexpr int NT = band_plane_threads<N>();
      constexpr size_t sm = band_plane_smem<T, N>();
      if ((e = group_attr(&k_plane_band_gc<T, N, NT>, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_band_gc<T, N, NT>), dim3(blocks), dim3(NT), sm, s, g, n);
      return hipGetLastError();
    }
  }
  if constexpr (plane_smem<T, N, N>() <= 160 * 1024) {
    if (kind == GK_LDS) {
      constexpr size_t sm = plane_smem<T, N, N>();
      if ((e = group_attr(&k_plane_gc<T, N, kPlaneThreads>, sm)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_plane_gc<T, N, kPlaneThreads>), dim3(blocks), dim3(kPlaneThreads), sm, s, g, n);
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_group_fft(int pass, int key, const GroupHpsi<T>* g, int n, int blocks, size_t smem, hipStream_t s) {
  if (n < 1 || blocks < 1) return hipErrorInvalidValue;
  const int kind = key / kGroupKeyN, len = key % kGroupKeyN;
  if (kind == GK_RT) return launch_group_fft_rt<T>(pass, g, n, blocks, smem, s);
  switch (len) {
#define PWH_CASE(N) \
  case N: return launch_group_n<T, N>(pass, kind, g, n, blocks, s);
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
}

template int group_fft_key<double>(int, const SphereMaps&, int, int*, size_t*, int*);
template int group_fft_key<float>(int, const SphereMaps&, int, int*, size_t*, int*);
template hipError_t launch_group_fft<double>(int, int, const GroupHpsi<double>*, int, int, size_t, hipStream_t);
template hipError_t launch_group_fft<float>(int, int, const GroupHpsi<float>*, int, int, size_t, hipStream_t);

template <int N>
static hipError_t launch_plane_rho_n(const SphereMaps& m, const void* tw1, const void* tw2, const cx<double>* W,
                                     int nb, int groups, const double* wts, double* part, hipStream_t s) {
  static const bool band_on = [] {
    const char* e = getenv("PWH_PLANE_BAND");
    return !(e && e[0] == '0');
  }();
  if constexpr (N % 4 == 0 && band_plane_smem<double, N>() <= 160 * 1024) {
    if (m.band_xy && band_on) {
      constexpr int NT = band_plane_threads<N>();
      constexpr size_t bsm = band_plane_smem<double, N>();
      static bool battr = false;
      if (!battr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_rho_band<N, NT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bsm);
        if (e != hipSuccess) return e;
        battr = true;
      }
      hipLaunchKernelGGL((k_plane_rho_band<N, NT>), dim3(m.nr3, groups), dim3(NT), bsm, s, m,
                         static_cast<const cx<double>*>(tw1), W, nb, groups, wts, part);
      return hipGetLastError();
    }
  }
  constexpr size_t smem = plane_smem<double, N, N>();
  if constexpr (smem > 160 * 1024) {
    return hipErrorInvalidValue;
  } else {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_rho<N, N>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return e;
      attr = true;
    }
    hipLaunchKernelGGL((k_plane_rho<N, N>), dim3(m.nr3, groups), dim3(kRhoThreads), smem, s, m,
                       static_cast<const cx<double>*>(tw1), static_cast<const cx<double>*>(tw2), W, nb,
                       groups, wts, part);
    return hipGetLastError();
  }
}

hipError_t launch_plane_rho(const SphereMaps& m, const void* tw1, const void* tw2, const cx<double>* W, int nb,
                            int groups, const double* wts, double* part, double* rho, cx<double>* Y, int chunk,
                            hipStream_t s) {
  if (groups < 1) return hipErrorInvalidValue;
  if (m.path != FFT_COMPILED || !plane_compiled_fits(m, true))
    return launch_plane_rho_rt(m, tw1, tw2, W, nb, groups, wts, part, rho, Y, chunk, s);
  hipError_t e = hipErrorInvalidValue;
  switch (m.nr1) {
#define PWH_CASE(N) \
  case N: e = launch_plane_rho_n<N>(m, tw1, tw2, W, nb, groups, wts, part, s); break;
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  const size_t n = (size_t)m.nnr;
  hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, groups, n, rho);
  return hipGetLastError();
}

hipError_t launch_fft_lines(cx<double>* data, int n, int inner, int nlines, const cx<double>* tw,
                            int sign, double scale, hipStream_t s) {
  if (nlines < 1 || inner < 1) return hipErrorInvalidValue;
  if (!fft_size_supported(n)) {
    FftPlan p;
    if (!fft_plan(n, &p)) return hipErrorInvalidValue;
    return launch_fft_lines_rt(data, p, inner, nlines, tw, sign, scale, s);
  }
  dim3 grid((nlines + kColTile - 1) / kColTile);
  switch (n) {
#define PWH_CASE(N)                                                                                \
  case N:                                                                                          \
    if (sign > 0)                                                                                  \
      hipLaunchKernelGGL((k_fft_lines<double, N, +1>), grid, dim3(kColThreads), 0, s, data, inner, \
                         nlines, tw, scale);                                                       \
    else                                                                                           \
      hipLaunchKernelGGL((k_fft_lines<double, N, -1>), grid, dim3(kColThreads), 0, s, data, inner, \
                         nlines, tw, scale);                                                       \
    return hipGetLastError();
    PWH_FFT_SIZES(PWH_CASE)
#undef PWH_CASE
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_zinv<double>(const SphereMaps&, const void*, const cx<double>*, int,
                                        cx<double>*, int, hipStream_t);
template hipError_t launch_zinv<float>(const SphereMaps&, const void*, const cx<float>*, int,
                                       cx<float>*, int, hipStream_t);
template hipError_t launch_plane<double>(const SphereMaps&, const void*, const void*, const double*,
                                         cx<double>*, int, cx<double>*, int, hipStream_t, int);
template hipError_t launch_plane<float>(const SphereMaps&, const void*, const void*, const float*,
                                        cx<float>*, int, cx<float>*, int, hipStream_t, int);
template hipError_t launch_zfwd_epilogue<double>(const SphereMaps&, const void*, const cx<double>*,
                                                 const double*, const cx<double>*,
                                                 const cx<double>*, const cx<double>*, cx<double>*,
                                                 int, int, double, double, double, double,
                                                 hipStream_t);
template hipError_t launch_zfwd_epilogue<float>(const SphereMaps&, const void*, const cx<float>*,
                                                const float*, const cx<float>*, const cx<float>*,
                                                const cx<float>*, cx<float>*, int, int, float,
                                                float, float, float, hipStream_t);

}  // namespace pwh
