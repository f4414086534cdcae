// pwh_internal.h -- types shared by the HIP kernels and the host orchestration of
// libpwh (MI355X / gfx950 Kohn-Sham H|psi> and Chebyshev-filtered subspace iteration).
//
// Data layout in HBM (all per system, one k-point):
//   * "sphere order": the ngw plane waves are held in an internal order that groups
//     them by FFT column (i1,i2) and, inside a column, by i3.  perm[k] gives the
//     reference index ig (0-based) of internal slot k.  Every wave-function block is
//     band-major [nb][ngw_pad] (plane waves contiguous, ld = ngw_pad).
//   * W: the z-transformed columns between the passes, [nb][nr3][ncol_pad].
//   * V: the local potential on the grid, [nr3][nr2][nr1] (x fastest), like the
//     reference's rho argument.
//   * beta: the projectors EIGR(G,a) * PKG_A(l,G), [P][ngw_pad] (P = sum over species of
//     na * lm rows, row (a, l) of a species at a * lm + l), in FP64 and FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace pwh {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return mk<T>(s * a.x, s * a.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> conjg(cx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by (sign * i)
template <int SIGN, typename T>
__host__ __device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  return SIGN > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

enum Precision { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2, PREC_BF16X6 = 3 };

// A runtime-length FFT plan: the Stockham stages' radices in order (2, 3, 4, 5, 7, 8 have
// register butterflies; any other prime p runs as a direct length-p DFT per output).
constexpr int kFftMaxStages = 16;
constexpr int kFftMaxLen = 4096;
struct FftPlan {
  int n;
  int nst;
  int radix[kFftMaxStages];
};
// Factorises n (1 <= n <= kFftMaxLen): radix 8 while 8 divides the rest (4, 4 for 16), then
// 4, 2, 3, 5, 7 and the remaining primes in increasing order.  False when n is out of range.
bool fft_plan(int n, FftPlan* p);
// Alternative radix orders for the tuner: 0 as fft_plan; 1 radix 4 preferred over 8; 2 the
// stages of order 0 reversed.
bool fft_plan_variant(int n, int variant, FftPlan* p);

// Which implementation the FFT passes of a system use.
enum FftPath {
  FFT_AUTO = 0,      // compiled plans when they cover the grid, else runtime plans in LDS,
                     // else runtime plans with the plane split into three passes
  FFT_COMPILED = 1,  // compiled per-length kernels (2,3,5-smooth lengths 12..128, square planes)
  FFT_RUNTIME = 2,   // runtime plans, one (i3, band) plane per workgroup in LDS
  FFT_SPLIT = 3,     // runtime plans, plane pass as y / x / y line passes through HBM
};

// Device-side description of the sphere <-> grid maps used by the FFT passes.
struct SphereMaps {
  int nr1, nr2, nr3, nnr;
  int ngw, ngw_pad;
  int ncol, ncol_pad;   // columns (i1,i2) that hold sphere points; padded to the column tile
  int nxa;              // number of distinct i1 among the columns
  // 1 when nr1 == nr2, nr1 % 4 == 0 and every column has i1, i2 in the band
  // [0, n/4) U [n - n/4, n), i.e. signed index in [-n/4, n/4): true for grids sized
  // for the density cutoff (4 ecut), where the sphere radius is below n/4
  int band_xy;
  // largest number of sphere entries in one tile of 8 << t consecutive columns (t = 0..3)
  int tile_max[4];
  int* colstart;        // [ncol_pad+1] first internal slot of each column
  int* kcol;            // [ngw] column of each internal slot
  int* kz;              // [ngw] i3 of each internal slot
  int* colxy;           // [ncol] i1 + nr1*i2 of each column
  int* xa;              // [nxa] active i1 values
  // runtime plans and the maps of the split plane pass
  int path;             // FftPath actually used (never FFT_AUTO once the system exists)
  FftPlan plan[3];      // lengths nr1, nr2, nr3
  int* colxa;           // [ncol] index in xa of each column's i1
  int* xcols;           // [ncol] the columns ordered by (i1, i2)
  int* xcstart;         // [nxa + 1] first position in xcols of each active i1
  // band_xy grids: [(nr1/2) x (nr1/2)] column index of (band row, band column), -1 where no
  // column; band index b is position b for b < n/4 and b + n/2 above
  int* bandmap;
};

// Launchers (pwh_fft.hip)
template <typename T>
hipError_t launch_zinv(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W,
                       int nb, hipStream_t s);
// Y: the split pass's intermediate (plane_split_elems(m, chunk) elements, chunk =
// plane_split_chunk(...)); unused by the other paths (may be null there).  cap: at most that
// many workgroups of a compiled plane kernel per CU (0: no cap), so the nonlocal GEMMs of
// the second stream find room beside them.
template <typename T>
hipError_t launch_plane(const SphereMaps& m, const void* tw1, const void* tw2, const T* V, cx<T>* W,
                        int nb, cx<T>* Y, int chunk, hipStream_t s, int cap = 0);
template <typename T>
hipError_t launch_zfwd_epilogue(const SphereMaps& m, const void* tw3, const cx<T>* W, const T* kin,
                                const cx<T>* X, const cx<T>* Hnl, const cx<T>* Z, cx<T>* out,
                                int ld, int nb, T inv_nnr, T alpha, T beta, T gamma,
                                hipStream_t s);
bool fft_size_supported(int n);
bool plane_supported(int n1, int n2);
// Runtime-plan passes (pwh_fft_rt.hip).  plane_rt_fits: the (i1, i2) plane and its
// transforms fit one workgroup (LDS and registers) in this precision.
bool plane_rt_fits(int n1, int n2, bool fp64);
// Elements of the split plane pass's intermediate Y[nbc][nr3][nxa][nr2] for nbc bands,
// and the band chunk that keeps it within `budget_bytes`.
size_t plane_split_elems(const SphereMaps& m, int nbc);
// the plane pass (and the density pass, for fp64) of this precision takes the split path
bool plane_uses_split(const SphereMaps& m, bool fp64);
// the compiled per-length kernels cover every pass of this grid
bool compiled_plans_cover(int n1, int n2, int n3);
constexpr size_t kSplitBudget = size_t(1) << 30;   // bytes of the split pass's intermediate
int plane_split_chunk(const SphereMaps& m, int nb, size_t bytes_per_complex, size_t budget_bytes);
// One-workgroup FP64 Hermitian eigensolver for n <= 128 (pwh_eig.hip): A (column-major,
// lda, upper triangle read) -> eigenvectors, w ascending; info[0] != 0 on a non-finite value.
size_t heev_small_scratch_doubles();
constexpr int kHeevSmallMax = 128;
// One-workgroup inverse Cholesky factor (pwh_chol.hip): S (column-major, lda, upper
// triangle read as Hermitian positive definite) -> R^{-1} with S = R^H R, zeros below the
// diagonal; info[0] = 0, or k + 1 when the k-th pivot is not positive (S then undefined).
constexpr int kCholSmallMax = 128;
hipError_t launch_chol_inv_small(cx<double>* S, int lda, int n, int* info, hipStream_t s);
hipError_t launch_heev_small(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                             hipStream_t s);
// Householder reduction of the Hermitian n x n matrix A (column-major, both triangles read)
// to real tridiagonal T = Q^H A Q (pwh_trd.hip), 3 <= n <= trd_max_n(): d, e (n, n-1), tau
// (n-1), the reflectors kept in work (trd_work_elems(n) complex values).  A is not
// modified.  tridiag_backtransform: Y (n x n, column-major) <- Q Y with that work and tau.
size_t trd_work_elems(int n);
int trd_max_n();
hipError_t tridiag_householder(const cx<double>* A, int n, cx<double>* work, double* d, double* e, cx<double>* tau,
                               hipStream_t s);
hipError_t tridiag_backtransform(const cx<double>* work, const cx<double>* tau, cx<double>* Y, int n, hipStream_t s);
// the eigenvalues of the symmetric tridiagonal matrix (d, e), ascending, by multisection
hipError_t tridiag_eigenvalues(const double* d, const double* e, int n, double* w, hipStream_t s);
// Address, on the current device, of the solver scratch held in a device-resident array
// (heev_small_scratch_doubles() doubles): no allocation, shared by every call on the device.
hipError_t heev_small_static_scratch(double** ptr);
// The same solver in two halves on one scratch: the values half (tridiagonal reduction and
// multisection) leaves w complete; the vectors half (inverse iteration, multiplet
// orthonormalisation, back-transformation) then writes the eigenvectors into A and may set
// info.  Together they are launch_heev_small.
hipError_t launch_heev_small_values(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                                    hipStream_t s);
hipError_t launch_heev_small_vectors(cx<double>* A, int lda, int n, double* w, int* info, double* scratch,
                                     hipStream_t s);
// Density of a block in sphere order (FP64): W (after launch_zinv) -> rho[nr3][nr2][nr1] =
// sum_b wts[b] |psi_b(r)|^2, the bands split over `groups` partial grids part[groups][nnr].
hipError_t launch_plane_rho(const SphereMaps& m, const void* tw1, const void* tw2, const cx<double>* W, int nb,
                            int groups, const double* wts, double* part, double* rho, cx<double>* Y, int chunk,
                            hipStream_t s);
// Dense FP64 transform of length n along an axis of stride `inner` for every line of an
// [outer][n][inner] array, in place; tw[m] = exp(+2 pi i m / n); sign +1 inverse, -1 forward.
hipError_t launch_fft_lines(cx<double>* data, int n, int inner, int nlines, const cx<double>* tw,
                            int sign, double scale, hipStream_t s);

// Runtime-plan launchers (pwh_fft_rt.hip), used by the ones above when the compiled plans
// do not cover the grid (or the system's path asks for them).
template <typename T>
hipError_t launch_zinv_rt(const SphereMaps& m, const void* tw3, const cx<T>* X, int ldx, cx<T>* W, int nb,
                          hipStream_t s);
template <typename T>
hipError_t launch_zfwd_rt(const SphereMaps& m, const void* tw3, const cx<T>* W, const T* kin, const cx<T>* X,
                          const cx<T>* Hnl, const cx<T>* Z, cx<T>* out, int ld, int nb, T inv_nnr, T alpha,
                          T beta, T gamma, hipStream_t s);
template <typename T>
hipError_t launch_plane_rt(const SphereMaps& m, const void* tw1, const void* tw2, const T* V, cx<T>* W, int nb,
                           cx<T>* Y, int chunk, hipStream_t s);
hipError_t launch_plane_rho_rt(const SphereMaps& m, const void* tw1, const void* tw2, const cx<double>* W, int nb,
                               int groups, const double* wts, double* part, double* rho, cx<double>* Y, int chunk,
                               hipStream_t s);
hipError_t launch_fft_lines_rt(cx<double>* data, const FftPlan& p, int inner, int nlines, const cx<double>* tw,
                               int sign, double scale, hipStream_t s);

// Grouped FFT passes of H|psi> over the systems of a batch (pwh_fft_rt.hip): one launch per
// pass, one descriptor per system (device memory), systems in order of their first blocks.
template <typename T>
struct GroupHpsi {
  SphereMaps m;
  const cx<T>* tw1;
  const cx<T>* tw2;
  const cx<T>* tw3;
  const T* V;
  const T* kin;
  const cx<T>* X;     // the block H acts on (ld), also the beta X term of the epilogue
  cx<T>* W;           // the system's column workspace
  const cx<T>* Hnl;   // nonlocal part (null: none)
  const cx<T>* Z;     // gamma Z term (null: none)
  cx<T>* out;
  int ld, nb, ct;      // ct: column tile of the runtime-plan i3 passes
  T inv_nnr, alpha, beta, gamma;
  int blk0;           // first block of the system in its launch's grid
};
// The runtime-plan shape of a system in the grouped passes: false when its plane needs
// the split pass or more than 256 lanes.
template <typename T>
bool group_fft_shape(const SphereMaps& m, int nb, int* ct, int* zblocks, int* pblocks, size_t* zsmem,
                     size_t* psmem);
// pass 0: i3 inverse, 1: plane, 2: i3 forward with the epilogue.  group_fft_key: the kernel
// a system's pass runs (systems with equal keys share a launch), its blocks and LDS, and
// the runtime column tile; -1 when the pass cannot be grouped.  launch_group_fft: one
// launch of a key over n systems (blocks: their total, smem: their largest).
template <typename T>
int group_fft_key(int pass, const SphereMaps& m, int nb, int* blocks, size_t* smem, int* ct);
template <typename T>
hipError_t launch_group_fft(int pass, int key, const GroupHpsi<T>* g, int n, int blocks, size_t smem, hipStream_t s);
template <typename T>
hipError_t launch_group_fft_rt(int pass, const GroupHpsi<T>* g, int n, int blocks, size_t smem, hipStream_t s);

// GEMM launchers (pwh_gemm.hip)
struct BetaGen {
  // the projectors of one species, beta[(a,l)][k] = EIGR(k,a) * PKG_A(l,k), stored
  // row-major [rows = na * lm][ld] (row (a, l) at a * lm + l)
  const void* B;   // cx<T>
  int rows, lm, ld;
  int target = 0;  // split-K block target of the projection (0: the library default)
  // FP32 projections and expansions with beta generated on the fly from factorised phases
  // instead of read from B (fly != 0): EIGR(k, a) = EI1(m1) EI2(m2) EI3(m3) of atom slot
  // slot0 + a, tables ei [slot][3][ldt] (cx<float>); mi[k]: the three table indices of
  // plane wave k packed 10 bits each; pkg [lm][ld] (float) of the species
  int ks = 0;       // LDS stage depth of the GEMM (16, 32, or 64 for FP32; 0: the library default)
  int fly = 0;
  const void* ei = nullptr;
  const int* mi = nullptr;
  const float* pkg = nullptr;
  int ldt = 0, slot0 = 0;
};
// C[(a,l)][b] = wl[l] * <beta_(a,l)|X_b> (wl null: unscaled), row-major [rows][nb]
template <typename T>
hipError_t gemm_proj(const BetaGen& bg, const cx<T>* X, int ldx, int nb, int K, cx<T>* C,
                     cx<T>* part, size_t part_elems, int mode, hipStream_t s, const T* wl = nullptr);
template <typename T>
hipError_t gemm_expand(const BetaGen& bg, const cx<T>* Q, int nb, int K, cx<T>* out, int ldo,
                       bool accumulate, int mode, hipStream_t s);
// C[m][n] = sum_k conj(A[m][k]) B[n][k]  (A: [M][lda], B: [N][ldb]); FP64 result
hipError_t gemm_cn_f64(const cx<double>* A, int lda, int M, const cx<double>* B, int ldb, int N,
                       int K, cx<double>* C, int ldc, cx<double>* part, size_t part_elems,
                       hipStream_t s);
// out[n][k] = sum_j A[j][k] Q[j][n]  (A: [J][lda], Q: [J][ldq] row j, out: [N][ldo])
hipError_t gemm_nn_f64(const cx<double>* A, int lda, int J, const cx<double>* Q, int ldq, int N,
                       int K, cx<double>* out, int ldo, hipStream_t s);
// The Gram matrix C = A^H A (A: [N][lda] rows = vectors), column-major N x N; only its
// upper triangle (m <= n) is computed, the strict lower triangle is left undefined.
hipError_t gemm_gram_upper_f64(const cx<double>* A, int lda, int N, int K, cx<double>* C, cx<double>* part,
                               size_t part_elems, hipStream_t s);
// gemm_nn_f64 with Q upper triangular in column-major storage (Q(j, n) = Q[j + n ldq] = 0
// for j > n): each column tile reduces over the rows up to its last column only.
hipError_t gemm_nn_tri_f64(const cx<double>* A, int lda, int J, const cx<double>* Q, int ldq, int N, int K,
                           cx<double>* out, int ldo, hipStream_t s);
hipError_t mfma_selftest(int kind, double* max_err);

// Grouped projector GEMMs over the systems (and species) of a batch (pwh_gemm.hip), one
// launch per kind.  Each problem keeps the split-K shape the one-system path gives it
// (proj_split_shape with the system's target), so the sums are the same.
template <typename T>
struct GroupProj {
  const cx<T>* A;     // beta rows of the species, [M][lda]
  const cx<T>* B;     // the block, [N][ldb]
  cx<T>* part;        // S * M * N partial sums
  cx<T>* C;           // Q rows of the species, [M][ldc]
  const T* wl;        // WNL(l) of the species
  int lda, ldb, ldc, M, N, K, kchunk, S, lm;
  int blk0;           // first block in the projection grid
  int rblk0;          // first block in the reduction grid
};
template <typename T>
struct GroupExpand {
  const cx<T>* A;     // beta rows, [J][lda]
  const cx<T>* Q;     // [J][ldq]
  cx<T>* out;         // [N][ldo]
  int lda, ldq, ldo, J, N, K, accumulate;
  int blk0;           // first block in the expansion grid
};
// Split-K shape of the projection C[M][N] over K in precision T (the one-system rule):
// S partial sums over chunks of kchunk; blocks of its grid.  0 when the grouped kernels
// cannot reproduce the one-system launch (another stage depth is configured).
template <typename T>
int proj_split_shape(int M, int N, int K, int target, int* S, int* kchunk, int* blocks);
int expand_grid_blocks(int K, int N);
template <typename T>
hipError_t launch_group_proj(const GroupProj<T>* g, int n, int blocks, int rblocks, hipStream_t s);
template <typename T>
hipError_t launch_group_expand(const GroupExpand<T>* g, int n, int blocks, hipStream_t s);

// Elementwise / small kernels (pwh_blas.hip)
template <typename Tin, typename Tout>
hipError_t convert_cx(const cx<Tin>* in, cx<Tout>* out, size_t n, hipStream_t s);
// one array of a grouped conversion: n complex values, from block blk0 on (256 per block)
struct GroupConv {
  const void* in;
  void* out;
  size_t n;
  int blk0;
};
template <typename Tin, typename Tout>
hipError_t launch_group_convert(const GroupConv* g, int n, int blocks, hipStream_t s);
hipError_t permute_in_f64(const cx<double>* ref, int ldr, const int* perm, int ngw, int ngw_pad,
                          cx<double>* out, int nb, hipStream_t s);
hipError_t permute_out_f64(const cx<double>* in, int ngw_pad, const int* perm, int ngw,
                           cx<double>* ref, int ldr, int nb, hipStream_t s);

}  // namespace pwh
