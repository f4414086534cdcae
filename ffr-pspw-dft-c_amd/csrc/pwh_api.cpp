// pwh_api.cpp -- host orchestration and the C ABI of libpwh.
//
// One library stream carries every launch.  A pwh_system holds, in HBM, the sphere
// maps, the projectors beta = EIGR x PKG_A in both FP64 and FP32, the local potential, and
// resident work buffers that grow on demand; nothing is allocated per call once a
// system has seen its largest block.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pwh.h"
#include "pwh_blas.h"
#include "pwh_internal.h"
#include "pwh_scf.h"
#include "pwh_setup_dev.h"

using namespace pwh;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(PWH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)
#define RBCHK(expr)                                                                     \
  do {                                                                                  \
    rocblas_status s_ = (expr);                                                         \
    if (s_ != rocblas_status_success)                                                   \
      return fail(PWH_ERR_HIP, std::string(#expr) + ": rocblas status " + std::to_string((int)s_)); \
  } while (0)
#define CHK(expr)                 \
  do {                            \
    int r_ = (expr);              \
    if (r_ != PWH_OK) return r_;  \
  } while (0)

enum TimerId { TM_ZINV, TM_PLANE, TM_ZFWD, TM_PROJ, TM_EXPAND, TM_HPSI, TM_RHO_SUMS, TM_RHO_ZINV, TM_RHO_PLANE,
               TM_RHO, TM_ENERGY, TM_COUNT };
constexpr size_t kRocblasWorkspace = size_t(128) << 20;   // bytes
const char* kTimerNames[TM_COUNT] = {"zinv", "plane", "zfwd", "proj", "expand", "hpsi",
                                     "rho_sums", "rho_zinv", "rho_plane", "rho", "energy"};

struct Ctx {
  bool ready = false;
  int device = 0;
  // bumped whenever a work buffer moves or a launch shape is retuned: a captured batch
  // graph holds buffer addresses and grid sizes, and is captured again after a change
  unsigned long long layout_gen = 0;
  hipStream_t stream = nullptr;
  rocblas_handle rb = nullptr;
  bool timing = false;
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[TM_COUNT] = {0};
  int count[TM_COUNT] = {0};
  // fixed device workspace handed to the rocBLAS handle
  void* ws = nullptr;
  size_t ws_bytes = 0;
};
Ctx g;

hipEvent_t take_event() {
  if (!g.pool.empty()) {
    hipEvent_t e = g.pool.back();
    g.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

// Records HIP events around a launch on the library stream when timing is enabled.
struct Timed {
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr;
  Timed(int i, hipStream_t s) : id(i), st(s) {
    if (g.timing) {
      a = take_event();
      hipEventRecord(a, st);
    }
  }
  ~Timed() {
    if (g.timing && a) {
      hipEvent_t b = take_event();
      hipEventRecord(b, st);
      g.pending.push_back({id, a, b});
    }
  }
};

void resolve_timers() {
  for (auto& r : g.pending) {
    hipEventSynchronize(r.b);
    float ms = 0;
    hipEventElapsedTime(&ms, r.a, r.b);
    g.total_ms[r.id] += ms;
    g.count[r.id] += 1;
    g.pool.push_back(r.a);
    g.pool.push_back(r.b);
  }
  g.pending.clear();
}

int ensure_init() {
  if (g.ready) return PWH_OK;
  return pwh_init(0);
}

// Device buffer that grows on demand.  Growing waits for the device first so no queued
// kernel (on the library stream or a system's second stream) still reads the old
// allocation.
template <typename T>
struct Buf {
  T* p = nullptr;
  size_t cap = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  int ensure(size_t n) {
    if (n <= cap) return PWH_OK;
    if (p) {
      if (g.ready) hipDeviceSynchronize();
      hipFree(p);
    }
    p = nullptr;
    cap = 0;
    ++g.layout_gen;
    hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess) return fail(PWH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    hipMemsetAsync(p, 0, sizeof(T) * std::max<size_t>(n, 1), g.stream);
    cap = n;
    return PWH_OK;
  }
  void release() {
    if (p) {
      if (g.ready) hipDeviceSynchronize();
      hipFree(p);
    }
    p = nullptr;
    cap = 0;
  }
};

template <typename T>
int upload(Buf<T>& b, const T* h, size_t n) {
  CHK(b.ensure(n));
  HIPCHK(hipMemcpyAsync(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice, g.stream));
  return PWH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// The system
// ---------------------------------------------------------------------------
struct pwh_system {
  int nr[3] = {0, 0, 0};
  int nnr = 0, ngw = 0, ngw_pad = 0;
  double tpiba2 = 0;
  SphereMaps maps{};
  std::vector<int> perm;
  Buf<int> d_perm, d_colstart, d_kcol, d_kz, d_colxy, d_xa, d_colxa, d_xcols, d_xcstart, d_bandmap;
  int fft_request = FFT_AUTO;   // the FftPath asked for (pwh_fft_path)
  int nsp = 0, nat = 0;
  std::vector<int> na, lm, aoff, poff;   // poff: first projector row of each species
  std::vector<double> wl;   // [nsp*4]
  // constants, both precisions
  // projectors beta[(a,l)][k] = EIGR(k,a) PKG_A(l,k), [nproj][ngw_pad], species sp from row poff[sp]
  Buf<cx<double>> Bt64;
  Buf<double> kin64, wl64, V64;
  Buf<cx<float>> Bt32;
  Buf<float> kin32, wl32, V32;
  bool vset = false;
  Buf<cx<double>> tw64[3];
  Buf<cx<float>> tw32[3];
  // FP64 work
  Buf<cx<double>> X, HX, A64, B64, C64, W64, Hnl64, Q64, part64, stage, S, Vl;
  Buf<double> evals, rwork, dres;
  Buf<double> rho_part, rho, rho_w, rho_q, rho_k;   // density pass (pwh_eval_rho_psi)
  // block solvers (pwh_davidson, pwh_lobpcg): Q = [X | R] or [X | W | P] and H Q, one block
  // of scratch, the small projected matrices and per-band vectors
  Buf<cx<double>> dq, dhq, dtmp, dT, dG, dE, dT1, dt1;
  Buf<cx<double>> lw;    // two n x n work matrices of the series second pass of CholeskyQR2
  Buf<double> dmax;      // its max |G - I|
  Buf<double> dlam, dnrm, dpart, dprec, dbo;
  Buf<double> dTl;       // the Lanczos T on the device
  Buf<rocblas_int> info;
  // FP32 work
  Buf<cx<float>> X32, Y32, Y1_32, W32, Hnl32, Q32, part32;
  // intermediate of the split plane pass (grids whose planes do not fit one workgroup)
  Buf<cx<double>> Ysp64;
  Buf<cx<float>> Ysp32;
  int resident_nb = 0;
  // split-K block target of the projections per precision (PREC_FP64/FP32/BF16X3/BF16X6; 0: the
  // library default), set by pwh_autotune
  int tune[4] = {0, 0, 0, 0};
  // LDS stage depth of the projection [p][0] and expansion [p][1] GEMMs per precision (0:
  // the library default), and the radix-order variant of the runtime FFT plans
  int gks[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  int fftv = 0;
  // plane-pass workgroups per CU while the nonlocal GEMMs overlap it, per precision (0: no
  // cap), set by pwh_autotune
  int pcap[4] = {0, 0, 0, 0};
  // second stream for the nonlocal GEMMs, which overlap the FFT passes (fork/join events)
  hipStream_t nl_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // third stream: ChFSI runs the eigenvector half of a Rayleigh-Ritz solve on it while the
  // next filter runs on the library stream (fork/join events)
  hipStream_t eig_stream = nullptr;
  hipEvent_t ev_eig_fork = nullptr, ev_eig_done = nullptr;
  rocblas_handle rb_eig = nullptr;   // rocSOLVER calls of the eigenvector half, on eig_stream
  // projectors generated on the fly in the FP32 GEMMs (pwh_projector_mode): per-plane-wave
  // packed table indices, PKG_A in sphere order [sum lm][ngw_pad], the FP32 phase tables
  int beta_fly = 0, nax = 0, ldt = 0;
  std::vector<int> pko;              // first PKG_A row of each species
  Buf<int> mi;
  Buf<float> pkg32;
  Buf<cx<float>> ei32;
  std::vector<float> pkg_host;       // kept from creation until the mode is set
};

namespace {

template <typename T> struct Sel;
template <> struct Sel<double> {
  static cx<double>* Bt(pwh_system* s) { return s->Bt64.p; }
  static double* kin(pwh_system* s) { return s->kin64.p; }
  static double* wl(pwh_system* s) { return s->wl64.p; }
  static double* V(pwh_system* s) { return s->V64.p; }
  static void* tw(pwh_system* s, int d) { return s->tw64[d].p; }
  static Buf<cx<double>>& W(pwh_system* s) { return s->W64; }
  static Buf<cx<double>>& Hnl(pwh_system* s) { return s->Hnl64; }
  static Buf<cx<double>>& Q(pwh_system* s) { return s->Q64; }
  static Buf<cx<double>>& part(pwh_system* s) { return s->part64; }
  static Buf<cx<double>>& Ysp(pwh_system* s) { return s->Ysp64; }
};
template <> struct Sel<float> {
  static cx<float>* Bt(pwh_system* s) { return s->Bt32.p; }
  static float* kin(pwh_system* s) { return s->kin32.p; }
  static float* wl(pwh_system* s) { return s->wl32.p; }
  static float* V(pwh_system* s) { return s->V32.p; }
  static void* tw(pwh_system* s, int d) { return s->tw32[d].p; }
  static Buf<cx<float>>& W(pwh_system* s) { return s->W32; }
  static Buf<cx<float>>& Hnl(pwh_system* s) { return s->Hnl32; }
  static Buf<cx<float>>& Q(pwh_system* s) { return s->Q32; }
  static Buf<cx<float>>& part(pwh_system* s) { return s->part32; }
  static Buf<cx<float>>& Ysp(pwh_system* s) { return s->Ysp32; }
};

// Band chunk of the split plane pass for nb bands in this precision (0: the plane pass of
// this precision does not split), with its intermediate ensured.
template <typename T>
int split_prepare(pwh_system* s, int nb, int* chunk) {
  *chunk = 0;
  if (!plane_uses_split(s->maps, std::is_same<T, double>::value)) return PWH_OK;
  *chunk = plane_split_chunk(s->maps, nb, sizeof(cx<T>), kSplitBudget);
  return Sel<T>::Ysp(s).ensure(plane_split_elems(s->maps, *chunk));
}

// out = alpha * H X + beta * X + gamma * Z  on a block of nb bands in sphere order.
int overlap_prepare(pwh_system* s) {
  if (s->nl_stream) return PWH_OK;
  HIPCHK(hipStreamCreateWithFlags(&s->nl_stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
  return PWH_OK;
}

// The fork/join costs two event operations per H|psi>; it pays only when the block is
// large enough that the GEMMs run concurrently with a still-busy FFT pass (Si64: 13133 PW
// x 128 bands gains 12%; 24 small cells of <= 5449 PW x 76 bands lose 25%).  Branches of
// a captured batch already run side by side, so they never fork.  With per-kernel timing
// on, everything stays on one stream so each event pair brackets one kernel running alone.
constexpr size_t kOverlapMinElems = size_t(1) << 19;
bool overlap_enabled(const pwh_system* s, int nb, hipStream_t st) {
  static const int force = [] {
    const char* e = getenv("PWH_NO_OVERLAP");
    return (e && e[0] == '1') ? 0 : 1;
  }();
  return force && !g.timing && st == g.stream && (size_t)nb * (size_t)s->ngw >= kOverlapMinElems;
}

// Ensures every work buffer hpsi<T> needs for nb bands (so a later call allocates nothing).
template <typename T>
int hpsi_prepare(pwh_system* s, int nb) {
  const int ld = s->ngw_pad;
  CHK(overlap_prepare(s));
  CHK(Sel<T>::W(s).ensure((size_t)nb * s->nr[2] * s->maps.ncol_pad));
  int chunk;
  CHK(split_prepare<T>(s, nb, &chunk));
  for (int sp = 0; sp < s->nsp; ++sp) {
    if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
    const int M = s->na[sp] * s->lm[sp];
    CHK(Sel<T>::Hnl(s).ensure((size_t)nb * ld));
    const size_t nproj = (size_t)s->poff.back() + (size_t)s->na.back() * s->lm.back();
    CHK(Sel<T>::Q(s).ensure(nproj * nb));
    size_t pe = 0;
    for (int m = 0; m < 4; ++m)
      if ((m == PREC_FP64) == std::is_same<T, double>::value)
        pe = std::max(pe, proj_part_elems(M, nb, s->ngw, s->tune[m]));
    CHK(Sel<T>::part(s).ensure(pe));
  }
  return PWH_OK;
}

// the on-the-fly projector fields of species sp (FP32 GEMMs only read them)
void beta_fly_fields(const pwh_system* s, int sp, BetaGen& bg) {
  if (!s->beta_fly) return;
  bg.fly = 1;
  bg.ei = s->ei32.p;
  bg.mi = s->mi.p;
  bg.pkg = s->pkg32.p + (size_t)s->pko[sp] * s->ngw_pad;
  bg.ldt = s->ldt;
  bg.slot0 = sp * s->nax;
}

template <typename T>
int hpsi(pwh_system* s, const cx<T>* X, cx<T>* out, int nb, double alpha, double beta,
         double gamma, const cx<T>* Z, int mode, hipStream_t st = nullptr) {
  if (!s->vset) return fail(PWH_ERR_ARG, "local potential not set (pwh_set_vloc)");
  if (!st) st = g.stream;
  Timed th(TM_HPSI, st);
  const int ld = s->ngw_pad;
  CHK(Sel<T>::W(s).ensure((size_t)nb * s->nr[2] * s->maps.ncol_pad));
  int chunk;
  CHK(split_prepare<T>(s, nb, &chunk));
  bool any_nl = false;
  for (int sp = 0; sp < s->nsp; ++sp) any_nl = any_nl || (s->na[sp] > 0 && s->lm[sp] > 0);
  // the nonlocal GEMMs run on a second stream, overlapping the FFT passes (running the
  // projections first and overlapping only the expansions measured no better)
  hipStream_t sn = st;
  const bool overlap = any_nl && overlap_enabled(s, nb, st);
  if (overlap) {
    CHK(overlap_prepare(s));
    HIPCHK(hipEventRecord(s->ev_fork, st));
    HIPCHK(hipStreamWaitEvent(s->nl_stream, s->ev_fork, 0));
    sn = s->nl_stream;
  }
  for (int sp = 0; sp < s->nsp; ++sp) {
    if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
    const int M = s->na[sp] * s->lm[sp];
    CHK(Sel<T>::Hnl(s).ensure((size_t)nb * ld));
    const size_t nproj = (size_t)s->poff.back() + (size_t)s->na.back() * s->lm.back();
    CHK(Sel<T>::Q(s).ensure(nproj * nb));
    const int target = s->tune[std::is_same<T, double>::value ? PREC_FP64
                               : (mode == PREC_BF16X3 || mode == PREC_BF16X6) ? mode : PREC_FP32];
    const size_t pe = proj_part_elems(M, nb, s->ngw, target);
    CHK(Sel<T>::part(s).ensure(pe));
    BetaGen bg{Sel<T>::Bt(s) + (size_t)s->poff[sp] * ld, M, s->lm[sp], ld, target};
    beta_fly_fields(s, sp, bg);
    bg.ks = s->gks[std::is_same<T, double>::value ? PREC_FP64 : mode][0];
    hipStream_t sp_st = sn;
    {
      // <beta|psi> scaled by WNL straight into Q by the split-K reduction
      Timed t(TM_PROJ, sp_st);
      HIPCHK(gemm_proj<T>(bg, X, ld, nb, s->ngw, Sel<T>::Q(s).p + (size_t)s->poff[sp] * nb, Sel<T>::part(s).p,
                          pe, mode, sp_st, Sel<T>::wl(s) + 4 * sp));
    }
  }
  bool have_nl = false;
  for (int sp = 0; sp < s->nsp; ++sp) {
    if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
    const int M = s->na[sp] * s->lm[sp];
    BetaGen bg{Sel<T>::Bt(s) + (size_t)s->poff[sp] * ld, M, s->lm[sp], ld};
    beta_fly_fields(s, sp, bg);
    bg.ks = s->gks[std::is_same<T, double>::value ? PREC_FP64 : mode][1];
    {
      Timed t(TM_EXPAND, sn);
      HIPCHK(gemm_expand<T>(bg, Sel<T>::Q(s).p + (size_t)s->poff[sp] * nb, nb, s->ngw, Sel<T>::Hnl(s).p, ld,
                            have_nl, mode, sn));
    }
    have_nl = true;
  }
  if (overlap) HIPCHK(hipEventRecord(s->ev_join, sn));
  {
    Timed t(TM_ZINV, st);
    HIPCHK(launch_zinv<T>(s->maps, Sel<T>::tw(s, 2), X, ld, Sel<T>::W(s).p, nb, st));
  }
  {
    Timed t(TM_PLANE, st);
    const int cap = overlap ? s->pcap[std::is_same<T, double>::value ? PREC_FP64 : mode] : 0;
    HIPCHK(launch_plane<T>(s->maps, Sel<T>::tw(s, 0), Sel<T>::tw(s, 1), Sel<T>::V(s), Sel<T>::W(s).p,
                           nb, Sel<T>::Ysp(s).p, chunk, st, cap));
  }
  if (overlap) HIPCHK(hipStreamWaitEvent(st, s->ev_join, 0));
  {
    Timed t(TM_ZFWD, st);
    HIPCHK(launch_zfwd_epilogue<T>(s->maps, Sel<T>::tw(s, 2), Sel<T>::W(s).p, Sel<T>::kin(s), X,
                                   have_nl ? Sel<T>::Hnl(s).p : nullptr, Z, out, ld, nb,
                                   (T)(1.0 / s->nnr), (T)alpha, (T)beta, (T)gamma, st));
  }
  return PWH_OK;
}

// chebyfilt.cpp:6-74 on device buffers b0 (= X in), b1, b2 (work); *result -> filtered block
template <typename T>
int chebyfilt_dev(pwh_system* s, cx<T>* b0, cx<T>* b1, cx<T>* b2, int nb, int degree, double lb,
                  double ub, int mode, cx<T>** result, hipStream_t st = nullptr, cx<T>** previous = nullptr,
                  double* sigma_out = nullptr, bool first_in_b1 = false) {
  const double e = (ub - lb) / 2.0;
  const double c = (ub + lb) / 2.0;
  double sigma = e / (lb - ub);
  const double sigma1 = sigma;
  cx<T>* X = b0;
  cx<T>* Y = b1;
  cx<T>* Y1 = b2;
  // first iterate (H X - c X) sigma1 / e, unless the caller formed it in b1 already
  if (!first_in_b1) CHK(hpsi<T>(s, X, Y, nb, sigma1 / e, -c * sigma1 / e, 0.0, nullptr, mode, st));
  for (int id = 2; id <= degree; ++id) {
    const double sigma2 = 1.0 / (2.0 / sigma1 - sigma);
    CHK(hpsi<T>(s, Y, Y1, nb, 2.0 * sigma2 / e, -c * 2.0 * sigma2 / e, -sigma * sigma2, X, mode, st));
    cx<T>* t = X;
    X = Y;
    Y = Y1;
    Y1 = t;
    sigma = sigma2;
  }
  *result = Y;
  if (previous) *previous = X;
  if (sigma_out) *sigma_out = sigma;
  return PWH_OK;
}

// The mixed-precision filters run their last recurrence step in FP64: rounding injected
// by the FP32 steps is damped by the steps after it, except in the last one, which would
// leave an FP32-sized error in the filtered block (it showed as a ~3e-7 Ha wander of the
// converged SCF total energy).  prev/cur: FP32 iterates m-2 and m-1; out: FP64 result.
// PWH_FILTER_LAST64=0 keeps every step in the filter's precision (for measurement).
bool filter_last_fp64() {
  static const bool on = [] {
    const char* e = getenv("PWH_FILTER_LAST64");
    return !(e && e[0] == '0');
  }();
  return on;
}

int cheby_last_step_fp64(pwh_system* s, const cx<float>* prev, const cx<float>* cur, cx<double>* out, int nb,
                         double lb, double ub, double sigma, hipStream_t st) {
  const size_t n = (size_t)nb * s->ngw_pad;
  const double e = (ub - lb) / 2.0, c = (ub + lb) / 2.0;
  const double sigma1 = e / (lb - ub);
  const double sigma2 = 1.0 / (2.0 / sigma1 - sigma);
  HIPCHK((convert_cx<float, double>(prev, s->A64.p, n, st)));
  HIPCHK((convert_cx<float, double>(cur, s->B64.p, n, st)));
  return hpsi<double>(s, s->B64.p, out, nb, 2.0 * sigma2 / e, -c * 2.0 * sigma2 / e, -sigma * sigma2, s->A64.p,
                      PREC_FP64, st);
}

int check_ld(pwh_system* s, int nb) {
  if (!s) return fail(PWH_ERR_ARG, "null system");
  if (nb < 1) return fail(PWH_ERR_ARG, "nstates < 1");
  return PWH_OK;
}

// X (ref order, host, [ngw x nb]) -> s->X (sphere order, FP64)
int load_block(pwh_system* s, const double* Xh, int nb) {
  const size_t nref = (size_t)s->ngw * nb;
  CHK(upload(s->stage, reinterpret_cast<const cx<double>*>(Xh), nref));
  CHK(s->X.ensure((size_t)nb * s->ngw_pad));
  HIPCHK(permute_in_f64(s->stage.p, s->ngw, s->d_perm.p, s->ngw, s->ngw_pad, s->X.p, nb, g.stream));
  return PWH_OK;
}

int store_block(pwh_system* s, const cx<double>* Xd, double* Xh, int nb) {
  const size_t nref = (size_t)s->ngw * nb;
  CHK(s->stage.ensure(nref));
  HIPCHK(permute_out_f64(Xd, s->ngw_pad, s->d_perm.p, s->ngw, s->stage.p, s->ngw, nb, g.stream));
  HIPCHK(hipMemcpyAsync(Xh, s->stage.p, sizeof(cx<double>) * nref, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

// H X in the given precision for the FP64 sphere-order block X -> out (FP64)
int apply_prec(pwh_system* s, const cx<double>* X, cx<double>* out, int nb, int precision) {
  if (precision == PWH_FP64) return hpsi<double>(s, X, out, nb, 1.0, 0.0, 0.0, nullptr, PREC_FP64);
  const size_t n = (size_t)nb * s->ngw_pad;
  CHK(s->X32.ensure(n));
  CHK(s->Y32.ensure(n));
  HIPCHK((convert_cx<double, float>(X, s->X32.p, n, g.stream)));
  CHK(hpsi<float>(s, s->X32.p, s->Y32.p, nb, 1.0, 0.0, 0.0, nullptr, precision));
  HIPCHK((convert_cx<float, double>(s->Y32.p, out, n, g.stream)));
  return PWH_OK;
}

// Chebyshev filter of the FP64 sphere-order block X in place.  HX: H X of this block in
// FP64 if the caller has it (the rotated H X of the Rayleigh-Ritz step just before, under
// the same potential); the first iterate is then formed from it without an H|psi>.
int filter_prec(pwh_system* s, cx<double>* X, int nb, int degree, double lb, double ub,
                int precision, const cx<double>* HX = nullptr) {
  const size_t n = (size_t)nb * s->ngw_pad;
  if (degree < 1) return PWH_OK;
  const double e = (ub - lb) / 2.0, c = (ub + lb) / 2.0, sigma1 = e / (lb - ub);
  CHK(s->A64.ensure(n));
  CHK(s->B64.ensure(n));
  if (precision == PWH_FP64) {
    if (HX) HIPCHK(axpby(HX, sigma1 / e, X, -c * sigma1 / e, s->A64.p, (int)n, g.stream));
    cx<double>* res = nullptr;
    CHK(chebyfilt_dev<double>(s, X, s->A64.p, s->B64.p, nb, degree, lb, ub, PREC_FP64, &res, g.stream,
                              nullptr, nullptr, HX != nullptr));
    if (res != X) HIPCHK(hipMemcpyAsync(X, res, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
    return PWH_OK;
  }
  CHK(s->X32.ensure(n));
  CHK(s->Y32.ensure(n));
  CHK(s->Y1_32.ensure(n));
  HIPCHK((convert_cx<double, float>(X, s->X32.p, n, g.stream)));
  if (HX) {
    HIPCHK(axpby(HX, sigma1 / e, X, -c * sigma1 / e, s->B64.p, (int)n, g.stream));
    HIPCHK((convert_cx<double, float>(s->B64.p, s->Y32.p, n, g.stream)));
  }
  cx<float>* res = nullptr;
  if (degree >= 2 && filter_last_fp64()) {
    cx<float>* prev = nullptr;
    double sigma = 0;
    CHK(chebyfilt_dev<float>(s, s->X32.p, s->Y32.p, s->Y1_32.p, nb, degree - 1, lb, ub, precision, &res,
                             g.stream, &prev, &sigma, HX != nullptr));
    return cheby_last_step_fp64(s, prev, res, X, nb, lb, ub, sigma, g.stream);
  }
  CHK(chebyfilt_dev<float>(s, s->X32.p, s->Y32.p, s->Y1_32.p, nb, degree, lb, ub, precision, &res, g.stream,
                           nullptr, nullptr, HX != nullptr));
  HIPCHK((convert_cx<float, double>(res, X, n, g.stream)));
  return PWH_OK;
}

// Ensures the buffers of filter_on_stream for nb bands in the given precision.
int filter_prepare(pwh_system* s, int nb, int precision) {
  const size_t n = (size_t)nb * s->ngw_pad;
  CHK(s->X.ensure(n));
  if (precision == PWH_FP64) {
    CHK(s->A64.ensure(n));
    CHK(s->B64.ensure(n));
    return hpsi_prepare<double>(s, nb);
  }
  CHK(s->X32.ensure(n));
  CHK(s->Y32.ensure(n));
  CHK(s->Y1_32.ensure(n));
  if (filter_last_fp64()) {
    CHK(s->A64.ensure(n));
    CHK(s->B64.ensure(n));
    CHK(hpsi_prepare<double>(s, nb));
  }
  return hpsi_prepare<float>(s, nb);
}

// Chebyshev filter of the resident FP64 block s->X on stream st; buffers must be
// prepared (filter_prepare), so this issues kernels and device copies only and can be
// captured into a graph.
int filter_on_stream(pwh_system* s, int nb, int degree, double lb, double ub, int precision,
                     hipStream_t st) {
  const size_t n = (size_t)nb * s->ngw_pad;
  if (precision == PWH_FP64) {
    cx<double>* res = nullptr;
    CHK(chebyfilt_dev<double>(s, s->X.p, s->A64.p, s->B64.p, nb, degree, lb, ub, PREC_FP64, &res, st));
    if (res != s->X.p)
      HIPCHK(hipMemcpyAsync(s->X.p, res, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, st));
    return PWH_OK;
  }
  HIPCHK((convert_cx<double, float>(s->X.p, s->X32.p, n, st)));
  cx<float>* res = nullptr;
  if (degree >= 2 && filter_last_fp64()) {
    cx<float>* prev = nullptr;
    double sigma = 0;
    CHK(chebyfilt_dev<float>(s, s->X32.p, s->Y32.p, s->Y1_32.p, nb, degree - 1, lb, ub, precision, &res, st,
                             &prev, &sigma));
    return cheby_last_step_fp64(s, prev, res, s->X.p, nb, lb, ub, sigma, st);
  }
  CHK(chebyfilt_dev<float>(s, s->X32.p, s->Y32.p, s->Y1_32.p, nb, degree, lb, ub, precision, &res, st));
  HIPCHK((convert_cx<float, double>(res, s->X.p, n, st)));
  return PWH_OK;
}

// PWH_CHOLQR_SERIES=0 keeps the Cholesky second pass in ChFSI; PWH_CHOLQR_SERIES_TOL sets
// the bound on n max|E| below which the series replaces it (default 1e-3; 0 never)
bool cholqr_series_on() {
  static const bool on = [] {
    const char* e = getenv("PWH_CHOLQR_SERIES");
    return !(e && e[0] == '0');
  }();
  return on;
}
double cholqr_series_tol() {
  static const double t = [] {
    const char* e = getenv("PWH_CHOLQR_SERIES_TOL");
    const double v = e ? atof(e) : 1e-3;
    return (v >= 0.0 && v <= 1e-3) ? v : 1e-3;
  }();
  return t;
}

// C = A B for column-major n x n device matrices (rocBLAS on the library stream)
int zgemm_nn_small(int n, const cx<double>* A, const cx<double>* B, cx<double>* C) {
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  RBCHK(rocblas_zgemm(g.rb, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                      reinterpret_cast<const rocblas_double_complex*>(A), n,
                      reinterpret_cast<const rocblas_double_complex*>(B), n, &zero,
                      reinterpret_cast<rocblas_double_complex*>(C), n));
  return PWH_OK;
}

// CholeskyQR2 (FP64) of the nb columns of X [nb][ld] (nrows valid rows), using tmp [nb][ld];
// dinfo: 2 ints of device scratch.  For nb <= 128 the factor inverse comes from the
// one-workgroup kernel and both passes are checked with one synchronisation at the end;
// above, from rocSOLVER's zpotrf + ztrtri on the upper triangle of the Gram matrix (the
// only half computed), with X R^-1 reducing over the nonzero half of R^-1 only.
// span_only (a caller that needs an orthonormal basis of span(X), not the R of a QR: the
// ChFSI loop, whose Rayleigh-Ritz step rotates the basis anyway), with lw (2 nb^2) and
// dmax (1) given: after the first pass G = X^H X = I + E is nearly the identity, and when
// nb max|e_ij| <= 1e-3 (so ||E||_2 <= ||E||_F <= 1e-3) the second pass is X (I + E)^{-1/2}
// by the series I - E/2 + 3E^2/8 - 5E^3/16 + 35E^4/128 (truncation 63/256 ||E||^5 / (1 -
// ||E||) <= 2.5e-16), three rocBLAS n x n products in place of the one-workgroup inverse
// Cholesky factor.
int cholqr2(cx<double>* X, int ld, int nrows, int nb, cx<double>* tmp, cx<double>* S,
            cx<double>* part, size_t part_elems, rocblas_int* dinfo, cx<double>* lw = nullptr,
            double* dmax = nullptr) {
  if (nb <= kCholSmallMax && lw && dmax) {
    HIPCHK(gemm_cn_f64(X, ld, nb, X, ld, nb, nrows, S, nb, part, part_elems, g.stream));
    HIPCHK(launch_chol_inv_small(S, nb, nb, reinterpret_cast<int*>(dinfo), g.stream));
    HIPCHK(gemm_nn_f64(X, ld, nb, S, nb, nb, nrows, tmp, ld, g.stream));
    HIPCHK(gemm_cn_f64(tmp, ld, nb, tmp, ld, nb, nrows, S, nb, part, part_elems, g.stream));
    HIPCHK(eye_residual(S, nb, dmax, g.stream));
    rocblas_int hinfo = 0;
    double emax = 0.0;
    HIPCHK(hipMemcpyAsync(&hinfo, dinfo, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipMemcpyAsync(&emax, dmax, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (hinfo != 0)
      return fail(PWH_ERR_LAPACK, "CholeskyQR: overlap matrix not positive definite (zpotrf info " +
                                      std::to_string(hinfo) + ")");
    if ((double)nb * emax <= cholqr_series_tol() && cholqr_series_tol() > 0.0) {
      cx<double>* A = lw;
      cx<double>* B = lw + (size_t)nb * nb;
      HIPCHK(axpy_eye(A, S, 35.0 / 128.0, -5.0 / 16.0, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, A, B));
      HIPCHK(axpy_eye(B, B, 1.0, 3.0 / 8.0, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, B, A));
      HIPCHK(axpy_eye(A, A, 1.0, -0.5, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, A, B));
      HIPCHK(axpy_eye(B, B, 1.0, 1.0, nb, g.stream));
      HIPCHK(gemm_nn_f64(tmp, ld, nb, B, nb, nb, nrows, X, ld, g.stream));
      return PWH_OK;
    }
    // not close enough to the identity: the Cholesky second pass on G = E + I
    HIPCHK(axpy_eye(S, S, 1.0, 1.0, nb, g.stream));
    HIPCHK(launch_chol_inv_small(S, nb, nb, reinterpret_cast<int*>(dinfo) + 1, g.stream));
    HIPCHK(gemm_nn_f64(tmp, ld, nb, S, nb, nb, nrows, X, ld, g.stream));
    HIPCHK(hipMemcpyAsync(&hinfo, dinfo + 1, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (hinfo != 0)
      return fail(PWH_ERR_LAPACK, "CholeskyQR: overlap matrix not positive definite (zpotrf info " +
                                      std::to_string(hinfo) + ")");
    return PWH_OK;
  }
  if (nb <= kCholSmallMax) {
    for (int pass = 0; pass < 2; ++pass) {
      cx<double>* src = (pass == 0) ? X : tmp;
      cx<double>* dst = (pass == 0) ? tmp : X;
      HIPCHK(gemm_cn_f64(src, ld, nb, src, ld, nb, nrows, S, nb, part, part_elems, g.stream));
      HIPCHK(launch_chol_inv_small(S, nb, nb, reinterpret_cast<int*>(dinfo) + pass, g.stream));
      HIPCHK(gemm_nn_f64(src, ld, nb, S, nb, nb, nrows, dst, ld, g.stream));
    }
    rocblas_int hinfo[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(hinfo, dinfo, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    for (int pass = 0; pass < 2; ++pass)
      if (hinfo[pass] != 0)
        return fail(PWH_ERR_LAPACK, "CholeskyQR: overlap matrix not positive definite (zpotrf info " +
                                        std::to_string(hinfo[pass]) + ")");
    return PWH_OK;
  }
  // nb > 128: rocSOLVER's zpotrf and ztrtri on the Gram matrix, of which only the upper
  // triangle is computed (all zpotrf reads); X R^-1 skips the zero half of R^-1.  With
  // span_only the second pass is the series above when the first left X^H X close to I.
  auto lapack_pass = [&](cx<double>* src, cx<double>* dst, bool have_gram) -> int {
    if (!have_gram) HIPCHK(gemm_gram_upper_f64(src, ld, nb, nrows, S, part, part_elems, g.stream));
    RBCHK(rocsolver_zpotrf(g.rb, rocblas_fill_upper, nb, reinterpret_cast<rocblas_double_complex*>(S),
                           nb, dinfo));
    rocblas_int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, dinfo, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (hinfo != 0)
      return fail(PWH_ERR_LAPACK, "CholeskyQR: overlap matrix not positive definite (zpotrf info " +
                                      std::to_string(hinfo) + ")");
    RBCHK(rocsolver_ztrtri(g.rb, rocblas_fill_upper, rocblas_diagonal_non_unit, nb,
                           reinterpret_cast<rocblas_double_complex*>(S), nb, dinfo));
    HIPCHK(zero_lower(S, nb, nb, g.stream));
    HIPCHK(gemm_nn_tri_f64(src, ld, nb, S, nb, nb, nrows, dst, ld, g.stream));
    return PWH_OK;
  };
  CHK(lapack_pass(X, tmp, false));
  if (lw && dmax && cholqr_series_tol() > 0.0) {
    HIPCHK(gemm_gram_upper_f64(tmp, ld, nb, nrows, S, part, part_elems, g.stream));
    HIPCHK(herm_fill_lower(S, nb, nb, g.stream));
    HIPCHK(eye_residual(S, nb, dmax, g.stream));
    double emax = 0.0;
    HIPCHK(hipMemcpyAsync(&emax, dmax, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if ((double)nb * emax <= cholqr_series_tol()) {
      cx<double>* A = lw;
      cx<double>* B = lw + (size_t)nb * nb;
      HIPCHK(axpy_eye(A, S, 35.0 / 128.0, -5.0 / 16.0, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, A, B));
      HIPCHK(axpy_eye(B, B, 1.0, 3.0 / 8.0, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, B, A));
      HIPCHK(axpy_eye(A, A, 1.0, -0.5, nb, g.stream));
      CHK(zgemm_nn_small(nb, S, A, B));
      HIPCHK(axpy_eye(B, B, 1.0, 1.0, nb, g.stream));
      HIPCHK(gemm_nn_f64(tmp, ld, nb, B, nb, nb, nrows, X, ld, g.stream));
      return PWH_OK;
    }
    HIPCHK(axpy_eye(S, S, 1.0, 1.0, nb, g.stream));   // G = E + I
    return lapack_pass(tmp, X, true);
  }
  return lapack_pass(tmp, X, false);
}

int sys_cholqr2(pwh_system* s, cx<double>* X, int nb, bool span_only = false) {
  const size_t n = (size_t)nb * s->ngw_pad;
  CHK(s->A64.ensure(n));
  CHK(s->S.ensure((size_t)nb * nb));
  const size_t pe = std::max(proj_part_elems(nb, nb, s->ngw), gram_part_elems(nb, s->ngw));
  CHK(s->part64.ensure(pe));
  CHK(s->info.ensure(2));
  const bool series = span_only && cholqr_series_on();
  if (series) {
    CHK(s->lw.ensure((size_t)2 * nb * nb));
    CHK(s->dmax.ensure(1));
  }
  return cholqr2(X, s->ngw_pad, s->ngw, nb, s->A64.p, s->S.p, s->part64.p, pe, s->info.p,
                 series ? s->lw.p : nullptr, series ? s->dmax.p : nullptr);
}

// FP64 Hermitian eigensolve on the GPU, eig_zheevd.cpp:5-34 ('V', 'U'): A (n x n,
// column-major) is replaced by its eigenvectors, w gets the eigenvalues ascending.
// Default: the one-workgroup solver of pwh_eig.hip for n <= 128 (Householder tridiagonal
// reduction, multisection, inverse iteration, back-transformation), rocSOLVER's divide and
// conquer above.  PWH_EIGSOLVER=evd|ev|evj|evdj forces rocSOLVER's divide and conquer, QR
// iteration, Jacobi or Jacobi with divide and conquer.  If the small solver meets a
// non-finite value, the saved input is solved again by rocSOLVER (zheevd).
// work: n doubles; info: 2 ints of device scratch.  Synchronises the library stream when
// the small solver ran (to read its flag).
// scratch of the small solver and the saved input for its fallback, kept for the life of
// the process
Buf<double>& heev_ws() {
  static Buf<double>* b = new Buf<double>();
  return *b;
}
Buf<cx<double>>& heev_save() {
  static Buf<cx<double>>* b = new Buf<cx<double>>();
  return *b;
}

bool small_eigensolver(int n) {
  const char* e = getenv("PWH_EIGSOLVER");
  return n <= kHeevSmallMax && (!e || !strcmp(e, "own"));
}

// Scratch of the mid-size path: the reduction's work, the reflectors, tau, d and e
Buf<cx<double>>& trd_ws() {
  static Buf<cx<double>>* b = new Buf<cx<double>>();
  return *b;
}
Buf<double>& trd_de() {
  static Buf<double>* b = new Buf<double>();
  return *b;
}

// 128 < n <= trd_max_n(): the multi-workgroup Householder reduction and back-transformation
// (pwh_trd.hip) around rocSOLVER's zstedc on the tridiagonal matrix.  A: Hermitian, upper
// triangle authoritative.  (rocSOLVER's zunmtr returned non-finite vectors for reflectors
// with tau = 0, as from a diagonal A; the back-transformation here skips those.)
int hermitian_eig_mid(cx<double>* A, int n, double* w, rocblas_int* info) {
  const size_t wk = trd_work_elems(n);
  CHK(trd_ws().ensure(wk + n));
  CHK(trd_de().ensure(2 * (size_t)n));
  cx<double>* work = trd_ws().p;
  cx<double>* tau = work + wk;
  double* d = trd_de().p;
  double* e = d + n;
  HIPCHK(herm_fill_lower(A, n, n, g.stream));
  HIPCHK(tridiag_householder(A, n, work, d, e, tau, g.stream));
  RBCHK(rocsolver_zstedc(g.rb, rocblas_evect_tridiagonal, n, d, e, reinterpret_cast<rocblas_double_complex*>(A), n,
                         info));
  HIPCHK(tridiag_backtransform(work, tau, A, n, g.stream));
  HIPCHK(hipMemcpyAsync(w, d, sizeof(double) * n, hipMemcpyDeviceToDevice, g.stream));
  return PWH_OK;
}

int hermitian_eig_impl(cx<double>* A, int n, double* w, double* work, rocblas_int* info) {
  const int mode = [] {
    const char* e = getenv("PWH_EIGSOLVER");   // read per call: a caller may switch solvers
    if (!e || !strcmp(e, "own")) return 0;
    if (!strcmp(e, "evj")) return 2;
    if (!strcmp(e, "ev")) return 3;
    if (!strcmp(e, "evdj")) return 4;
    return 1;
  }();
  auto* a = reinterpret_cast<rocblas_double_complex*>(A);
  if (mode == 0 && n <= kHeevSmallMax) {
    Buf<double>* ws = &heev_ws();
    Buf<cx<double>>* save = &heev_save();
    CHK(ws->ensure(heev_small_scratch_doubles()));
    CHK(save->ensure((size_t)n * n));
    HIPCHK(hipMemcpyAsync(save->p, A, sizeof(cx<double>) * n * n, hipMemcpyDeviceToDevice, g.stream));
    HIPCHK(launch_heev_small(A, n, n, w, reinterpret_cast<int*>(info), ws->p, g.stream));
    rocblas_int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, info, sizeof(flag), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (flag == 0) return PWH_OK;
    HIPCHK(hipMemcpyAsync(A, save->p, sizeof(cx<double>) * n * n, hipMemcpyDeviceToDevice, g.stream));
    RBCHK(rocsolver_zheevd(g.rb, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, work, info));
    return PWH_OK;
  }
  if (mode == 0 && n <= trd_max_n()) return hermitian_eig_mid(A, n, w, info);
  if (mode == 2) {
    RBCHK(rocsolver_zheevj(g.rb, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, n,
                           a, n, 0.0, work, 100, info + 1, w, info));
  } else if (mode == 3) {
    RBCHK(rocsolver_zheev(g.rb, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, work, info));
  } else if (mode == 4) {
    RBCHK(rocsolver_zheevdj(g.rb, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, info));
  } else {
    RBCHK(rocsolver_zheevd(g.rb, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, work, info));
  }
  return PWH_OK;
}

// Rayleigh-Ritz (ChebySCFLoop.cpp:292-306) on the orthonormal block X: X <- X Q, HX <- (H X) Q,
// evals (device) ascending.
int rayleigh_ritz_dev(pwh_system* s, cx<double>* X, int nb) {
  const int ld = s->ngw_pad;
  const size_t n = (size_t)nb * ld;
  CHK(s->HX.ensure(n));
  CHK(s->B64.ensure(n));
  CHK(s->S.ensure((size_t)nb * nb));
  CHK(s->evals.ensure(nb));
  CHK(s->rwork.ensure(nb));
  CHK(s->info.ensure(2));
  const size_t pe = proj_part_elems(nb, nb, s->ngw);
  CHK(s->part64.ensure(pe));
  CHK(hpsi<double>(s, X, s->HX.p, nb, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  HIPCHK(gemm_cn_f64(X, ld, nb, s->HX.p, ld, nb, s->ngw, s->S.p, nb, s->part64.p, pe, g.stream));
  HIPCHK(symmetrize(s->S.p, nb, nb, g.stream));
  CHK(hermitian_eig_impl(s->S.p, nb, s->evals.p, s->rwork.p, s->info.p));
  rocblas_int hinfo = 0;
  HIPCHK(hipMemcpyAsync(&hinfo, s->info.p, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (hinfo != 0) return fail(PWH_ERR_LAPACK, "zheevd info " + std::to_string(hinfo));
  HIPCHK(gemm_nn_f64(X, ld, nb, s->S.p, nb, nb, s->ngw, s->B64.p, ld, g.stream));
  HIPCHK(hipMemcpyAsync(X, s->B64.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  HIPCHK(gemm_nn_f64(s->HX.p, ld, nb, s->S.p, nb, nb, s->ngw, s->B64.p, ld, g.stream));
  HIPCHK(hipMemcpyAsync(s->HX.p, s->B64.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  return PWH_OK;
}

// The Rayleigh-Ritz step of rayleigh_ritz_dev in two halves, for the ChFSI loop: the
// eigenvector half runs on the system's third stream while the next filter runs on the
// library stream.  The filter is linear, p(H)(X Q) = (p(H) X) Q, so the next filter takes
// the unrotated X with its H X, and Q is applied to the filtered block afterwards.
int eig_overlap_prepare(pwh_system* s) {
  if (s->eig_stream) return PWH_OK;
  HIPCHK(hipStreamCreateWithFlags(&s->eig_stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&s->ev_eig_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&s->ev_eig_done, hipEventDisableTiming));
  RBCHK(rocblas_create_handle(&s->rb_eig));
  RBCHK(rocblas_set_stream(s->rb_eig, s->eig_stream));
  return PWH_OK;
}

bool mid_eigensolver(int n) {
  const char* e = getenv("PWH_EIGSOLVER");
  return n > kHeevSmallMax && n <= trd_max_n() && (!e || !strcmp(e, "own"));
}

bool chfsi_overlap_on(int nb) {
  static const bool on = [] {
    const char* e = getenv("PWH_CHFSI_OVERLAP");
    return !(e && e[0] == '0');
  }();
  return on && (small_eigensolver(nb) || mid_eigensolver(nb));
}

// The two halves of the mid-size solver: the values half reduces S to tridiagonal form
// (mid_values_half) and, once the vectors half is launched, takes the eigenvalues by
// multisection (mid_values_sterf: into evals; info[1] cleared); the vectors half, on the
// third stream, runs zstedc on copies of the tridiagonal matrix into S (info into info[0])
// and the back-transformation.  (rocSOLVER's dsterf took ~60 ms at n = 432.)
int mid_values_half(pwh_system* s, int nb) {
  const size_t wk = trd_work_elems(nb);
  CHK(trd_ws().ensure(wk + nb));
  CHK(trd_de().ensure(4 * (size_t)nb));
  cx<double>* work = trd_ws().p;
  cx<double>* tau = work + wk;
  double* d = trd_de().p;
  double* e = d + nb;
  HIPCHK(tridiag_householder(s->S.p, nb, work, d, e, tau, g.stream));
  double* d2 = e + nb;
  double* e2 = d2 + nb;
  HIPCHK(hipMemcpyAsync(s->evals.p, d, sizeof(double) * nb, hipMemcpyDeviceToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(e2, e, sizeof(double) * nb, hipMemcpyDeviceToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(d2, d, sizeof(double) * nb, hipMemcpyDeviceToDevice, g.stream));
  return PWH_OK;
}

int mid_values_sterf(pwh_system* s, int nb) {
  const double* d = trd_de().p;
  const double* e2 = trd_de().p + 3 * (size_t)nb;
  HIPCHK(tridiag_eigenvalues(d, e2, nb, s->evals.p, g.stream));
  HIPCHK(hipMemsetAsync(s->info.p + 1, 0, sizeof(rocblas_int), g.stream));
  return PWH_OK;
}

int mid_vectors_half(pwh_system* s, int nb) {
  const size_t wk = trd_work_elems(nb);
  cx<double>* work = trd_ws().p;
  cx<double>* tau = work + wk;
  double* d2 = trd_de().p + 2 * (size_t)nb;
  double* e = trd_de().p + nb;
  RBCHK(rocsolver_zstedc(s->rb_eig, rocblas_evect_tridiagonal, nb, d2, e,
                         reinterpret_cast<rocblas_double_complex*>(s->S.p), nb, s->info.p));
  HIPCHK(tridiag_backtransform(work, tau, s->S.p, nb, s->eig_stream));
  return PWH_OK;
}

// First half (ChebySCFLoop.cpp:292-300): HX = H X, S = sym(X^H HX), the eigenvalues of S
// (ascending) into s->evals and eig_host; then the eigenvector half is launched on the third
// stream.  rr_finish joins it.
int rr_begin(pwh_system* s, cx<double>* X, int nb, double* eig_host) {
  const int ld = s->ngw_pad;
  const size_t n = (size_t)nb * ld;
  CHK(s->HX.ensure(n));
  CHK(s->B64.ensure(n));
  CHK(s->S.ensure((size_t)nb * nb));
  CHK(s->evals.ensure(nb));
  CHK(s->rwork.ensure(nb));
  CHK(s->info.ensure(2));
  CHK(heev_ws().ensure(heev_small_scratch_doubles()));
  CHK(heev_save().ensure((size_t)nb * nb));
  CHK(eig_overlap_prepare(s));
  const size_t pe = proj_part_elems(nb, nb, s->ngw);
  CHK(s->part64.ensure(pe));
  CHK(hpsi<double>(s, X, s->HX.p, nb, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  HIPCHK(gemm_cn_f64(X, ld, nb, s->HX.p, ld, nb, s->ngw, s->S.p, nb, s->part64.p, pe, g.stream));
  HIPCHK(symmetrize(s->S.p, nb, nb, g.stream));
  HIPCHK(hipMemcpyAsync(heev_save().p, s->S.p, sizeof(cx<double>) * nb * nb, hipMemcpyDeviceToDevice, g.stream));
  int* info = reinterpret_cast<int*>(s->info.p);
  const bool mid = nb > kHeevSmallMax;
  if (mid) CHK(mid_values_half(s, nb));
  else HIPCHK(launch_heev_small_values(s->S.p, nb, nb, s->evals.p, info, heev_ws().p, g.stream));
  HIPCHK(hipEventRecord(s->ev_eig_fork, g.stream));
  HIPCHK(hipStreamWaitEvent(s->eig_stream, s->ev_eig_fork, 0));
  if (mid) CHK(mid_vectors_half(s, nb));
  else HIPCHK(launch_heev_small_vectors(s->S.p, nb, nb, s->evals.p, info, heev_ws().p, s->eig_stream));
  HIPCHK(hipEventRecord(s->ev_eig_done, s->eig_stream));
  if (mid) CHK(mid_values_sterf(s, nb));
  HIPCHK(hipMemcpyAsync(eig_host, s->evals.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g.stream));
  rocblas_int vinfo = 0;
  if (mid) HIPCHK(hipMemcpyAsync(&vinfo, s->info.p + 1, sizeof(vinfo), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (vinfo != 0) {
    // dsterf did not converge: join the vectors half and solve the saved matrix with zheevd
    HIPCHK(hipStreamSynchronize(s->eig_stream));
    HIPCHK(hipMemcpyAsync(s->S.p, heev_save().p, sizeof(cx<double>) * nb * nb, hipMemcpyDeviceToDevice, g.stream));
    RBCHK(rocsolver_zheevd(g.rb, rocblas_evect_original, rocblas_fill_upper, nb,
                           reinterpret_cast<rocblas_double_complex*>(s->S.p), nb, s->evals.p, s->rwork.p,
                           s->info.p));
    HIPCHK(hipMemcpyAsync(eig_host, s->evals.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipEventRecord(s->ev_eig_done, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
  }
  return PWH_OK;
}

// Second half: waits for the eigenvectors (the small solver's flag sends the saved matrix to
// rocSOLVER's zheevd, as hermitian_eig_impl does, and refreshes eig_host), then X <- X Q and,
// with rotate_hx, HX <- HX Q.
int rr_finish(pwh_system* s, cx<double>* X, int nb, bool rotate_hx, double* eig_host) {
  const int ld = s->ngw_pad;
  const size_t n = (size_t)nb * ld;
  HIPCHK(hipStreamWaitEvent(g.stream, s->ev_eig_done, 0));
  rocblas_int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, s->info.p, sizeof(flag), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (flag != 0) {
    HIPCHK(hipMemcpyAsync(s->S.p, heev_save().p, sizeof(cx<double>) * nb * nb, hipMemcpyDeviceToDevice,
                          g.stream));
    RBCHK(rocsolver_zheevd(g.rb, rocblas_evect_original, rocblas_fill_upper, nb,
                           reinterpret_cast<rocblas_double_complex*>(s->S.p), nb, s->evals.p, s->rwork.p,
                           s->info.p));
    rocblas_int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, s->info.p, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipMemcpyAsync(eig_host, s->evals.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (hinfo != 0) return fail(PWH_ERR_LAPACK, "zheevd info " + std::to_string(hinfo));
  }
  HIPCHK(gemm_nn_f64(X, ld, nb, s->S.p, nb, nb, s->ngw, s->B64.p, ld, g.stream));
  HIPCHK(hipMemcpyAsync(X, s->B64.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  if (rotate_hx) {
    HIPCHK(gemm_nn_f64(s->HX.p, ld, nb, s->S.p, nb, nb, s->ngw, s->B64.p, ld, g.stream));
    HIPCHK(hipMemcpyAsync(s->HX.p, s->B64.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  }
  return PWH_OK;
}

// Residual report (ChebySCFLoop.cpp:415-433) for X: HX, G = sym(X^H HX), R = HX - X G.
// hx_fresh: s->HX already holds H X for this X (after rayleigh_ritz_dev, same potential)
int residuals_dev(pwh_system* s, const cx<double>* X, int nb, double* res_host, bool hx_fresh = false) {
  const int ld = s->ngw_pad;
  const size_t n = (size_t)nb * ld;
  CHK(s->HX.ensure(n));
  CHK(s->B64.ensure(n));
  CHK(s->S.ensure((size_t)nb * nb));
  CHK(s->dres.ensure(nb));
  const size_t pe = proj_part_elems(nb, nb, s->ngw);
  CHK(s->part64.ensure(pe));
  if (!hx_fresh) CHK(hpsi<double>(s, X, s->HX.p, nb, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  HIPCHK(gemm_cn_f64(X, ld, nb, s->HX.p, ld, nb, s->ngw, s->S.p, nb, s->part64.p, pe, g.stream));
  HIPCHK(symmetrize(s->S.p, nb, nb, g.stream));
  HIPCHK(gemm_nn_f64(X, ld, nb, s->S.p, nb, nb, s->ngw, s->B64.p, ld, g.stream));
  HIPCHK(resid_norms(s->HX.p, s->B64.p, nullptr, ld, s->ngw, nb, s->dres.p, g.stream));
  HIPCHK(hipMemcpyAsync(res_host, s->dres.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

// lanczos.cpp:8-101 on the sphere-order vector v0 (device, normalised in place).
// T (host, ldT x k column-major), fnorm2 = ||f||^2, f (device) the final residual.
// lanczos.cpp:8-101 on device vectors, enqueued on the library stream with every scalar of
// the recurrence (norms, the Gram-Schmidt coefficients h and s, T in s->dTl, ld = k) on the
// device: no host round trip between the k steps.
constexpr int kLanczosDotSlices = 16;
int lanczos_enqueue(pwh_system* s, cx<double>* v0, int k, cx<double>* f) {
  const int ld = s->ngw_pad, n = s->ngw;
  cx<double>* V = s->Vl.p;
  cx<double>* Hv = s->C64.p;
  cx<double>* dh = s->S.p;
  cx<double>* ds = s->S.p + k;
  cx<double>* dn = s->S.p + 2 * k;
  cx<double>* dpart = s->S.p + 2 * k + 2;      // k * kLanczosDotSlices partial sums
  double* dT = s->dTl.p;
  auto multi_dotc = [&](const cx<double>* A, int lda, int j, const cx<double>* x, int nn, cx<double>* out,
                        hipStream_t st) {
    return multi_dotc_split(A, lda, j, x, nn, dpart, kLanczosDotSlices, out, st);
  };
  HIPCHK(fill_zero(dT, (size_t)k * k, g.stream));
  // v0 <- v0 / ||v0|| (the reference scales the caller's vector in place), V(:,1) = v0
  HIPCHK(multi_dotc(v0, ld, 1, v0, n, dn, g.stream));
  HIPCHK(scale_rnorm(v0, dn, v0, n, nullptr, g.stream));
  HIPCHK(axpby(v0, 1.0, nullptr, 0.0, V, n, g.stream));
  CHK(hpsi<double>(s, V, Hv, 1, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  // h_1 = v^H H v, T(1,1) = Re h_1; f = Hv - h_1 v; s_1 = v^H f; f -= s_1 v
  HIPCHK(multi_dotc(V, ld, 1, Hv, n, dh, g.stream));
  HIPCHK(tcol(dh, nullptr, 1, dT, g.stream));
  HIPCHK(multi_axpy_sub(V, ld, 1, dh, Hv, f, n, g.stream));
  HIPCHK(multi_dotc(V, ld, 1, f, n, ds, g.stream));
  HIPCHK(multi_axpy_sub(V, ld, 1, ds, f, f, n, g.stream));
  for (int j = 2; j <= k; ++j) {
    // T(j, j-1) = ||f||, V(:,j) = f / ||f||, Hv = H V(:,j)
    cx<double>* vj = V + (size_t)(j - 1) * ld;
    HIPCHK(multi_dotc(f, ld, 1, f, n, dn, g.stream));
    HIPCHK(scale_rnorm(f, dn, vj, n, dT + (j - 1) + (size_t)k * (j - 2), g.stream));
    CHK(hpsi<double>(s, vj, Hv, 1, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
    // h = V^H Hv, f = Hv - V h; one re-orthogonalisation: s = V^H f, f -= V s;
    // T(1:j, j) = Re(h + s)
    HIPCHK(multi_dotc(V, ld, j, Hv, n, dh, g.stream));
    HIPCHK(multi_axpy_sub(V, ld, j, dh, Hv, f, n, g.stream));
    HIPCHK(multi_dotc(V, ld, j, f, n, ds, g.stream));
    HIPCHK(multi_axpy_sub(V, ld, j, ds, f, f, n, g.stream));
    HIPCHK(tcol(dh, ds, j, dT + (size_t)k * (j - 1), g.stream));
  }
  HIPCHK(multi_dotc(f, ld, 1, f, n, dn, g.stream));
  return PWH_OK;
}

// T [ldT x k] (host) is filled as the reference fills it: T(i, j) = Re(h_i + s_i) for
// j >= 2, T(1, 1) = Re h_1 without the re-orthogonalisation correction, T(j, j-1) = ||f||;
// *fnorm2 = ||f||^2 of the final residual.
int lanczos_dev(pwh_system* s, cx<double>* v0, int k, double* T, int ldT, cx<double>* f,
                double* fnorm2) {
  const int ld = s->ngw_pad;
  CHK(s->Vl.ensure((size_t)k * ld));
  CHK(s->C64.ensure((size_t)ld));              // Hv
  CHK(s->S.ensure((size_t)(2 + kLanczosDotSlices) * k + 2));   // h, s, ||.||^2, dot partials
  CHK(s->dTl.ensure((size_t)k * k));
  CHK(lanczos_enqueue(s, v0, k, f));
  std::vector<double> Th((size_t)k * k);
  cx<double> nf;
  HIPCHK(hipMemcpyAsync(Th.data(), s->dTl.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(&nf, s->S.p + 2 * k, sizeof(nf), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  for (int i = 0; i < k * k; ++i) T[i] = 0.0;   // the reference clears k*k entries (lanczos.cpp:26)
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) T[i + (size_t)ldT * j] = Th[i + (size_t)k * j];
  *fnorm2 = nf.x;
  return PWH_OK;
}

// Eigenvalues of the symmetric matrix held in the upper triangle of T (k x k, ld ldT),
// dsyev_('N','U') of ChebySCFLoop.cpp:189, on the GPU.
int sym_eigvals_upper(const double* T, int ldT, int k, std::vector<double>& w) {
  Buf<double> dA, dD, dE;
  Buf<rocblas_int> di;
  std::vector<double> A((size_t)k * k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) A[i + (size_t)k * j] = T[i + (size_t)ldT * j];
  CHK(upload(dA, A.data(), A.size()));
  CHK(dD.ensure(k));
  CHK(dE.ensure(k));
  CHK(di.ensure(1));
  RBCHK(rocsolver_dsyevd(g.rb, rocblas_evect_none, rocblas_fill_upper, k, dA.p, k, dD.p, dE.p, di.p));
  w.resize(k);
  rocblas_int hinfo = 0;
  HIPCHK(hipMemcpyAsync(w.data(), dD.p, sizeof(double) * k, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(&hinfo, di.p, sizeof(hinfo), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  dA.release(); dD.release(); dE.release(); di.release();
  if (hinfo != 0) return fail(PWH_ERR_LAPACK, "dsyevd info " + std::to_string(hinfo));
  return PWH_OK;
}

// induced_norm_mat, ChebySCFLoop.cpp:7-34 (squared 2-norm of A * 1/sqrt(N)).
double induced_norm_sq(const double* A, int ldA, int N) {
  const double d = 1.0 / std::sqrt((double)N);
  double res = 0;
  for (int i = 0; i < N; ++i) {
    double v = 0;
    for (int j = 0; j < N; ++j) v += A[i + (size_t)ldA * j] * d;
    res += v * v;
  }
  return res;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device buffers of the system-free entry points (dense FFT, eigensolver), kept for the
// life of the process so repeated calls allocate nothing after the first.
struct Scratch {
  Buf<cx<double>> fft, fft_tw[3], eig_a;
  int fft_n[3] = {0, 0, 0};
  Buf<double> eig_w, eig_work;
  Buf<rocblas_int> eig_info;
};
Scratch& scratch() {
  static Scratch* s = new Scratch();   // never destroyed: no HIP calls after runtime teardown
  return *s;
}

std::vector<cx<double>> twiddles(int n) {
  std::vector<cx<double>> t(n);
  for (int m = 0; m < n; ++m) {
    const double a = 2.0 * M_PI * (double)m / (double)n;
    t[m] = mk<double>(std::cos(a), std::sin(a));
  }
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* pwh_last_error(void) { return g_err.c_str(); }

int pwh_device_count(int* count) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (count) *count = n;
  return PWH_OK;
}

int pwh_init(int device) {
  if (g.ready && g.device == device) return PWH_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(PWH_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(PWH_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  if (!g.stream) HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  if (!g.rb) {
    RBCHK(rocblas_create_handle(&g.rb));
    RBCHK(rocblas_set_stream(g.rb, g.stream));
    // one fixed device workspace for rocBLAS/rocSOLVER (zpotrf, ztrtri, zheevd up to
    // n = 432 need a few MiB): without it the handle re-allocates whenever consecutive
    // calls ask for different sizes, each time with a device synchronisation
    HIPCHK(hipMalloc(&g.ws, kRocblasWorkspace));
    g.ws_bytes = kRocblasWorkspace;
    RBCHK(rocblas_set_workspace(g.rb, g.ws, g.ws_bytes));
  }
  g.device = device;
  g.ready = true;
  return PWH_OK;
}

int pwh_selftest_mfma(int kind, double* max_err) {
  CHK(ensure_init());
  if (kind < 0 || kind > 2 || !max_err) return fail(PWH_ERR_ARG, "kind must be 0, 1 or 2");
  HIPCHK(mfma_selftest(kind, max_err));
  return PWH_OK;
}

int pwh_grid_supported(int nr1, int nr2, int nr3) {
  FftPlan p;
  return fft_plan(nr1, &p) && fft_plan(nr2, &p) && fft_plan(nr3, &p) ? 1 : 0;
}

int pwh_system_create(const pwh_system_desc* d, pwh_system** out) {
  if (!d || !out) return fail(PWH_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->nr1 < 1 || d->nr2 < 1 || d->nr3 < 1 || d->ngw < 1 || d->ngwx < d->ngw || !d->n123 ||
      !d->xkg || d->nsp < 0 || d->nax < 0 || d->nlmax < 1)
    return fail(PWH_ERR_ARG, "invalid system description");
  if (!pwh_grid_supported(d->nr1, d->nr2, d->nr3))
    return fail(PWH_ERR_UNSUPPORTED, "no FFT plan for grid " + std::to_string(d->nr1) + "x" +
                                         std::to_string(d->nr2) + "x" + std::to_string(d->nr3) +
                                         " (each length must be 2.." + std::to_string(kFftMaxLen) + ")");
  CHK(ensure_init());
  auto s = new pwh_system();
  auto bail = [&](int code) { pwh_system_destroy(s); return code; };
  s->nr[0] = d->nr1; s->nr[1] = d->nr2; s->nr[2] = d->nr3;
  s->nnr = d->nr1 * d->nr2 * d->nr3;
  s->ngw = d->ngw;
  s->ngw_pad = (d->ngw + 63) / 64 * 64;
  s->tpiba2 = d->tpiba2;

  // species / projectors: lm_end = L_MAX^2 - 2 L_LOC + 1 (ApplyHam_block.cpp:69-73)
  s->nsp = d->nsp;
  s->na.resize(d->nsp); s->lm.resize(d->nsp); s->aoff.resize(d->nsp); s->wl.assign(4 * d->nsp, 0.0);
  for (int sp = 0; sp < d->nsp; ++sp) {
    const int lm = d->l_max[sp] * d->l_max[sp] - 2 * d->l_loc[sp] + 1;
    if (lm > 4) { delete s; return fail(PWH_ERR_ARG, "lm_end > 4 is not implemented"); }
    if (d->na[sp] > d->nax) { delete s; return fail(PWH_ERR_ARG, "na > nax"); }
    s->lm[sp] = std::max(lm, 0);
    s->na[sp] = d->na[sp];
    s->aoff[sp] = s->nat;
    s->nat += d->na[sp];
    for (int l = 0; l < s->lm[sp]; ++l) s->wl[4 * sp + l] = d->wnl[sp + d->nsp * l];
  }

  // sphere packing: group plane waves by column (i1,i2), then by i3
  const int nr1 = d->nr1, nr2 = d->nr2, ngw = d->ngw;
  std::vector<long long> key(ngw);
  for (int ig = 0; ig < ngw; ++ig) {
    const unsigned n = d->n123[ig];
    if (n < 1 || (int)n > s->nnr) { delete s; return fail(PWH_ERR_ARG, "N123 entry out of range"); }
    const int idx = (int)n - 1;
    const int i1 = idx % nr1, i2 = (idx / nr1) % nr2, i3 = idx / (nr1 * nr2);
    key[ig] = ((long long)(i1 + nr1 * i2) << 20) | i3;
  }
  s->perm.resize(ngw);
  for (int i = 0; i < ngw; ++i) s->perm[i] = i;
  std::stable_sort(s->perm.begin(), s->perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  for (int k = 1; k < ngw; ++k)
    if (key[s->perm[k]] == key[s->perm[k - 1]]) { delete s; return fail(PWH_ERR_ARG, "duplicate N123 entries"); }
  std::vector<int> colstart, kcol(ngw), kz(ngw), colxy, xa;
  std::vector<char> xseen(nr1, 0);
  for (int k = 0; k < ngw; ++k) {
    const long long kk = key[s->perm[k]];
    const int xy = (int)(kk >> 20);
    if (k == 0 || xy != (int)(key[s->perm[k - 1]] >> 20)) {
      colstart.push_back(k);
      colxy.push_back(xy);
      if (!xseen[xy % nr1]) { xseen[xy % nr1] = 1; }
    }
    kcol[k] = (int)colxy.size() - 1;
    kz[k] = (int)(kk & ((1 << 20) - 1));
  }
  for (int x = 0; x < nr1; ++x) if (xseen[x]) xa.push_back(x);
  const int ncol = (int)colxy.size();
  const int ncol_pad = (ncol + 15) / 16 * 16;
  while ((int)colstart.size() < ncol_pad + 1) colstart.push_back(ngw);
  // the split plane pass's maps: each column's active-i1 index, the columns ordered by
  // (i1, i2) and where each active i1 starts in that order
  std::vector<int> xidx(nr1, -1), colxa(ncol), xcols(ncol), xcstart(xa.size() + 1, 0);
  for (size_t a = 0; a < xa.size(); ++a) xidx[xa[a]] = (int)a;
  for (int c = 0; c < ncol; ++c) {
    colxa[c] = xidx[colxy[c] % nr1];
    ++xcstart[colxa[c] + 1];
  }
  for (size_t a = 0; a < xa.size(); ++a) xcstart[a + 1] += xcstart[a];
  {
    std::vector<int> fill(xcstart.begin(), xcstart.end() - 1);
    for (int c = 0; c < ncol; ++c) xcols[fill[colxa[c]]++] = c;   // colxy ascends in i2 within an i1
  }
  int rc;
  if ((rc = upload(s->d_perm, s->perm.data(), ngw)) ||
      (rc = upload(s->d_colstart, colstart.data(), colstart.size())) ||
      (rc = upload(s->d_kcol, kcol.data(), ngw)) || (rc = upload(s->d_kz, kz.data(), ngw)) ||
      (rc = upload(s->d_colxy, colxy.data(), ncol)) || (rc = upload(s->d_xa, xa.data(), xa.size())) ||
      (rc = upload(s->d_colxa, colxa.data(), ncol)) || (rc = upload(s->d_xcols, xcols.data(), ncol)) ||
      (rc = upload(s->d_xcstart, xcstart.data(), xcstart.size())))
    return bail(rc);
  s->maps.nr1 = nr1; s->maps.nr2 = nr2; s->maps.nr3 = d->nr3; s->maps.nnr = s->nnr;
  s->maps.ngw = ngw; s->maps.ngw_pad = s->ngw_pad;
  s->maps.ncol = ncol; s->maps.ncol_pad = ncol_pad; s->maps.nxa = (int)xa.size();
  {
    auto in_band = [](int i, int n) { return i < n / 4 || i >= n - n / 4; };
    bool band = nr1 == nr2 && nr1 % 4 == 0;
    for (int c = 0; band && c < ncol; ++c)
      band = in_band(colxy[c] % nr1, nr1) && in_band(colxy[c] / nr1, nr2);
    s->maps.band_xy = band ? 1 : 0;
    s->maps.bandmap = nullptr;
    if (band) {
      const int h = nr1 / 2, q = nr1 / 4;
      std::vector<int> bmap((size_t)h * h, -1);
      auto band_index = [&](int i) { return i < q ? i : i - nr1 / 2; };
      for (int c = 0; c < ncol; ++c)
        bmap[(size_t)band_index(colxy[c] / nr1) * h + band_index(colxy[c] % nr1)] = c;
      if ((rc = upload(s->d_bandmap, bmap.data(), bmap.size()))) return bail(rc);
      s->maps.bandmap = s->d_bandmap.p;
    }
    for (int t = 0; t < 4; ++t) {
      const int ct = 8 << t;
      int mx = 0;
      for (int c0 = 0; c0 < ncol_pad; c0 += ct)
        mx = std::max(mx, colstart[std::min(c0 + ct, ncol_pad)] - colstart[c0]);
      s->maps.tile_max[t] = mx;
    }
  }
  s->maps.colstart = s->d_colstart.p; s->maps.kcol = s->d_kcol.p; s->maps.kz = s->d_kz.p;
  s->maps.colxy = s->d_colxy.p; s->maps.xa = s->d_xa.p;
  s->maps.colxa = s->d_colxa.p; s->maps.xcols = s->d_xcols.p; s->maps.xcstart = s->d_xcstart.p;
  for (int dd = 0; dd < 3; ++dd) fft_plan(s->nr[dd], &s->maps.plan[dd]);
  s->maps.path = compiled_plans_cover(nr1, nr2, d->nr3) ? FFT_COMPILED : FFT_RUNTIME;

  // constants in sphere order; the projectors beta = EIGR x PKG_A (ApplyHam_block.cpp:88-95)
  // are formed once in FP64 and rounded once for the FP32 copy
  const int ld = s->ngw_pad;
  s->poff.assign(s->nsp, 0);
  int nproj = 0;
  for (int sp = 0; sp < s->nsp; ++sp) { s->poff[sp] = nproj; nproj += s->na[sp] * s->lm[sp]; }
  std::vector<cx<double>> Bt((size_t)std::max(nproj, 1) * ld, mk<double>(0, 0));
  std::vector<double> kin(ld, 0.0);
  const cx<double>* eigr = reinterpret_cast<const cx<double>*>(d->eigr);
  for (int sp = 0; sp < s->nsp; ++sp)
    for (int ia = 0; ia < s->na[sp]; ++ia)
      for (int l = 0; l < s->lm[sp]; ++l) {
        cx<double>* row = Bt.data() + (size_t)(s->poff[sp] + ia * s->lm[sp] + l) * ld;
        for (int k = 0; k < ngw; ++k) {
          const size_t ig = s->perm[k];
          const cx<double> e = eigr[ig + (size_t)d->ngwx * (ia + (size_t)d->nax * sp)];
          const double pk = d->pkg_a[l + (size_t)d->nlmax * (ig + (size_t)d->ngwx * sp)];
          row[k] = mk<double>(e.x * pk, e.y * pk);
        }
      }
  for (int k = 0; k < ngw; ++k) kin[k] = 0.5 * d->xkg[s->perm[k]] * d->tpiba2;
  // PKG_A in sphere order and the packed phase-table indices of each plane wave, for the
  // projectors generated on the fly (pwh_projector_mode); the tables are built there
  {
    s->nax = d->nax;
    s->pko.assign(s->nsp, 0);
    int rows = 0;
    for (int sp = 0; sp < s->nsp; ++sp) { s->pko[sp] = rows; rows += s->lm[sp]; }
    s->pkg_host.assign((size_t)std::max(rows, 1) * ld, 0.f);
    for (int sp = 0; sp < s->nsp; ++sp)
      for (int l = 0; l < s->lm[sp]; ++l)
        for (int k = 0; k < ngw; ++k)
          s->pkg_host[(size_t)(s->pko[sp] + l) * ld + k] =
              (float)d->pkg_a[l + (size_t)d->nlmax * (s->perm[k] + (size_t)d->ngwx * sp)];
    s->ldt = std::max(nr1, std::max(nr2, d->nr3)) + 2;
    std::vector<int> mi(ld, 0);
    const int nrv[3] = {nr1, nr2, d->nr3};
    for (int k = 0; k < ngw; ++k) {
      unsigned g3 = d->n123[s->perm[k]] - 1;
      int packed = 0;
      for (int dd = 0; dd < 3; ++dd) {
        const int i = (int)(g3 % (unsigned)nrv[dd]);
        g3 /= (unsigned)nrv[dd];
        const int m = 2 * i > nrv[dd] ? i - nrv[dd] : i;   // the signed Miller index
        packed |= (m + (nrv[dd] + 1) / 2) << (10 * dd);
      }
      mi[k] = packed;
    }
    if ((rc = upload(s->mi, mi.data(), mi.size()))) return bail(rc);
  }
  std::vector<cx<float>> Bt32(Bt.size());
  std::vector<float> kin32(kin.size()), wl32(s->wl.size());
  for (size_t i = 0; i < Bt.size(); ++i) Bt32[i] = mk<float>((float)Bt[i].x, (float)Bt[i].y);
  for (size_t i = 0; i < kin.size(); ++i) kin32[i] = (float)kin[i];
  for (size_t i = 0; i < s->wl.size(); ++i) wl32[i] = (float)s->wl[i];
  std::vector<double> wlv = s->wl;
  if (wlv.empty()) wlv.assign(4, 0.0), wl32.assign(4, 0.f);
  if ((rc = upload(s->Bt64, Bt.data(), Bt.size())) ||
      (rc = upload(s->kin64, kin.data(), kin.size())) || (rc = upload(s->wl64, wlv.data(), wlv.size())) ||
      (rc = upload(s->Bt32, Bt32.data(), Bt32.size())) ||
      (rc = upload(s->kin32, kin32.data(), kin32.size())) || (rc = upload(s->wl32, wl32.data(), wl32.size())))
    return bail(rc);
  for (int dd = 0; dd < 3; ++dd) {
    auto t = twiddles(s->nr[dd]);
    if (dd == 0 && s->nr[0] == 64) {
      // the 64 x 64 plane pass's per-lane table [j][r] = exp(+2 pi i j r / 64) follows
      for (int j = 0; j < 8; ++j)
        for (int r = 0; r < 8; ++r) t.push_back(t[(j * r) % 64]);
    }
    if (dd == 0 && s->nr[0] == 96) {
      // the 96 x 96 plane pass's per-lane table [j][s] = exp(+2 pi i j s / 96) follows
      for (int j = 0; j < 8; ++j)
        for (int q = 0; q < 12; ++q) t.push_back(t[(j * q) % 96]);
    }
    std::vector<cx<float>> t32(t.size());
    for (size_t i = 0; i < t.size(); ++i) t32[i] = mk<float>((float)t[i].x, (float)t[i].y);
    if ((rc = upload(s->tw64[dd], t.data(), t.size())) || (rc = upload(s->tw32[dd], t32.data(), t32.size())))
      return bail(rc);
  }
  hipError_t e = hipStreamSynchronize(g.stream);
  if (e != hipSuccess) { pwh_system_destroy(s); return fail(PWH_ERR_HIP, hipGetErrorString(e)); }
  *out = s;
  return PWH_OK;
}

int pwh_system_destroy(pwh_system* s) {
  if (!s) return PWH_OK;
  if (g.stream) hipStreamSynchronize(g.stream);
  for (auto* b : {&s->d_perm, &s->d_colstart, &s->d_kcol, &s->d_kz, &s->d_colxy, &s->d_xa, &s->d_bandmap}) b->release();
  for (auto* b : {&s->Bt64, &s->X, &s->HX, &s->A64, &s->B64, &s->C64, &s->W64, &s->Hnl64,
                  &s->Q64, &s->part64, &s->stage, &s->S, &s->Vl})
    b->release();
  for (auto* b : {&s->kin64, &s->wl64, &s->V64, &s->evals, &s->rwork, &s->dres, &s->rho_part, &s->rho,
                  &s->rho_w, &s->rho_q, &s->rho_k})
    b->release();
  for (auto* b : {&s->Bt32, &s->X32, &s->Y32, &s->Y1_32, &s->W32, &s->Hnl32, &s->Q32, &s->part32})
    b->release();
  for (auto* b : {&s->kin32, &s->wl32, &s->V32}) b->release();
  for (int d = 0; d < 3; ++d) { s->tw64[d].release(); s->tw32[d].release(); }
  s->info.release();
  if (s->nl_stream) { hipStreamSynchronize(s->nl_stream); hipStreamDestroy(s->nl_stream); }
  if (s->ev_fork) hipEventDestroy(s->ev_fork);
  if (s->ev_join) hipEventDestroy(s->ev_join);
  if (s->eig_stream) { hipStreamSynchronize(s->eig_stream); hipStreamDestroy(s->eig_stream); }
  if (s->rb_eig) rocblas_destroy_handle(s->rb_eig);
  if (s->ev_eig_fork) hipEventDestroy(s->ev_eig_fork);
  if (s->ev_eig_done) hipEventDestroy(s->ev_eig_done);
  delete s;
  return PWH_OK;
}

int pwh_system_info(const pwh_system* s, int* ngw, int* ncol, int* nxa, size_t* bytes) {
  if (!s) return fail(PWH_ERR_ARG, "null system");
  if (ngw) *ngw = s->ngw;
  if (ncol) *ncol = s->maps.ncol;
  if (nxa) *nxa = s->maps.nxa;
  if (bytes) {
    size_t b = 0;
    auto add = [&](size_t cap, size_t el) { b += cap * el; };
    add(s->Bt64.cap, 16); add(s->Bt32.cap, 8);
    add(s->X.cap, 16); add(s->HX.cap, 16); add(s->A64.cap, 16); add(s->B64.cap, 16);
    add(s->W64.cap, 16); add(s->W32.cap, 8); add(s->Hnl64.cap, 16); add(s->Hnl32.cap, 8);
    add(s->X32.cap, 8); add(s->Y32.cap, 8); add(s->Y1_32.cap, 8); add(s->part64.cap, 16);
    add(s->part32.cap, 8); add(s->V64.cap, 8); add(s->V32.cap, 4); add(s->Vl.cap, 16);
    add(s->dq.cap, 16); add(s->dhq.cap, 16); add(s->dtmp.cap, 16);
    add(s->rho_part.cap, 8);
    *bytes = b;
  }
  return PWH_OK;
}

int pwh_fft_path(pwh_system* s, int path, int* used) {
  if (!s || path < FFT_AUTO || path > FFT_SPLIT) return fail(PWH_ERR_ARG, "bad fft_path arguments");
  const bool compiled = compiled_plans_cover(s->nr[0], s->nr[1], s->nr[2]);
  if (path == FFT_COMPILED && !compiled)
    return fail(PWH_ERR_UNSUPPORTED, "no compiled FFT plans for this grid (2,3,5-smooth 12..128, square planes)");
  const int eff = path == FFT_AUTO ? (compiled ? FFT_COMPILED : FFT_RUNTIME) : path;
  if (eff != s->maps.path) {
    if (g.stream) HIPCHK(hipStreamSynchronize(g.stream));
    s->maps.path = eff;
    ++g.layout_gen;   // captured batch graphs hold the old launch arguments
  }
  s->fft_request = path;
  if (used) *used = eff;
  return PWH_OK;
}

int pwh_set_vloc(pwh_system* s, const double* rho, int nnr) {
  if (!s || !rho) return fail(PWH_ERR_ARG, "null argument");
  if (nnr != s->nnr)
    return fail(PWH_ERR_ARG, "nnr != NNR : " + std::to_string(nnr) + " " + std::to_string(s->nnr));
  CHK(upload(s->V64, rho, nnr));
  std::vector<float> v32(rho, rho + nnr);
  CHK(upload(s->V32, v32.data(), v32.size()));
  HIPCHK(hipStreamSynchronize(g.stream));
  s->vset = true;
  return PWH_OK;
}

int pwh_apply_ham_block(pwh_system* s, const double* vin, double* vout, int ngw_ik, int nstates,
                        int precision) {
  CHK(check_ld(s, nstates));
  if (!vin || !vout) return fail(PWH_ERR_ARG, "null vector");
  if (ngw_ik != s->ngw) return fail(PWH_ERR_ARG, "NGW_ik does not match the system");
  if (precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  CHK(load_block(s, vin, nstates));
  CHK(s->HX.ensure((size_t)nstates * s->ngw_pad));
  CHK(apply_prec(s, s->X.p, s->HX.p, nstates, precision));
  return store_block(s, s->HX.p, vout, nstates);
}

int pwh_chebyfilt(pwh_system* s, double* X, int nstates, int degree, double lb, double ub,
                  int precision) {
  CHK(check_ld(s, nstates));
  if (!X) return fail(PWH_ERR_ARG, "null X");
  if (precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  if (!(ub > lb)) return fail(PWH_ERR_ARG, "chebyfilt needs ub > lb");
  CHK(load_block(s, X, nstates));
  CHK(filter_prec(s, s->X.p, nstates, degree, lb, ub, precision));
  return store_block(s, s->X.p, X, nstates);
}

int pwh_ortho_qr(double* X, int nbasis, int nstates) {
  if (!X || nbasis < 1 || nstates < 1 || nstates > nbasis) return fail(PWH_ERR_ARG, "bad ortho_qr arguments");
  CHK(ensure_init());
  const size_t n = (size_t)nbasis * nstates;
  Buf<cx<double>> dX, dT, dS, dP;
  Buf<rocblas_int> di;
  const size_t pe = proj_part_elems(nstates, nstates, nbasis);
  int rc;
  if ((rc = upload(dX, reinterpret_cast<const cx<double>*>(X), n)) || (rc = dT.ensure(n)) ||
      (rc = dS.ensure((size_t)nstates * nstates)) || (rc = dP.ensure(pe)) || (rc = di.ensure(2))) {
    dX.release(); dT.release(); dS.release(); dP.release(); di.release();
    return rc;
  }
  rc = cholqr2(dX.p, nbasis, nbasis, nstates, dT.p, dS.p, dP.p, pe, di.p);
  if (rc == PWH_ERR_LAPACK) {
    // X^H X is not numerically positive definite (cond(X) above ~1e8, or rank deficient):
    // Householder QR and the explicit Q, as the reference does (ortho_qr.cpp:19-26), on the
    // untouched input.  dS (nstates^2 >= nstates) holds tau.
    rc = upload(dX, reinterpret_cast<const cx<double>*>(X), n);
    auto* a = reinterpret_cast<rocblas_double_complex*>(dX.p);
    auto* tau = reinterpret_cast<rocblas_double_complex*>(dS.p);
    if (rc == PWH_OK && rocsolver_zgeqrf(g.rb, nbasis, nstates, a, nbasis, tau) != rocblas_status_success)
      rc = fail(PWH_ERR_LAPACK, "rocsolver_zgeqrf failed");
    if (rc == PWH_OK && rocsolver_zungqr(g.rb, nbasis, nstates, nstates, a, nbasis, tau) != rocblas_status_success)
      rc = fail(PWH_ERR_LAPACK, "rocsolver_zungqr failed");
  }
  if (rc == PWH_OK) {
    hipError_t e = hipMemcpyAsync(X, dX.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) rc = fail(PWH_ERR_HIP, hipGetErrorString(e));
  }
  hipStreamSynchronize(g.stream);
  dX.release(); dT.release(); dS.release(); dP.release(); di.release();
  return rc;
}

int pwh_fft3d(double* data, int nr1, int nr2, int nr3, int inverse) {
  if (!data || nr1 < 1 || nr2 < 1 || nr3 < 1) return fail(PWH_ERR_ARG, "bad fft3d arguments");
  const int nr[3] = {nr1, nr2, nr3};
  if (!pwh_grid_supported(nr1, nr2, nr3))
    return fail(PWH_ERR_UNSUPPORTED, "no FFT plan for a length outside 2.." + std::to_string(kFftMaxLen));
  CHK(ensure_init());
  Scratch& w = scratch();
  const size_t nnr = (size_t)nr1 * nr2 * nr3;
  for (int d = 0; d < 3; ++d) {
    if (w.fft_n[d] != nr[d]) {
      const auto t = twiddles(nr[d]);
      CHK(upload(w.fft_tw[d], t.data(), t.size()));
      w.fft_n[d] = nr[d];
    }
  }
  CHK(upload(w.fft, reinterpret_cast<const cx<double>*>(data), nnr));
  const int sign = inverse ? +1 : -1;
  const double scale = inverse ? 1.0 : 1.0 / (double)nnr;
  // z lines (stride nr1*nr2), y lines (stride nr1), x lines (contiguous, scaled)
  HIPCHK(launch_fft_lines(w.fft.p, nr3, nr1 * nr2, nr1 * nr2, w.fft_tw[2].p, sign, 1.0, g.stream));
  HIPCHK(launch_fft_lines(w.fft.p, nr2, nr1, nr1 * nr3, w.fft_tw[1].p, sign, 1.0, g.stream));
  HIPCHK(launch_fft_lines(w.fft.p, nr1, 1, nr2 * nr3, w.fft_tw[0].p, sign, scale, g.stream));
  HIPCHK(hipMemcpyAsync(data, w.fft.p, sizeof(cx<double>) * nnr, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_eig_zheevd(double* A, int ldA, double* lambda, int n) {
  if (!A || !lambda || n < 1 || ldA < n) return fail(PWH_ERR_ARG, "bad eig_zheevd arguments");
  CHK(ensure_init());
  Scratch& w = scratch();
  CHK(w.eig_a.ensure((size_t)n * n));
  CHK(w.eig_w.ensure(n));
  CHK(w.eig_work.ensure(n));
  CHK(w.eig_info.ensure(2));
  HIPCHK(hipMemcpy2DAsync(w.eig_a.p, sizeof(cx<double>) * n, A, sizeof(cx<double>) * ldA,
                          sizeof(cx<double>) * n, n, hipMemcpyHostToDevice, g.stream));
  CHK(hermitian_eig_impl(w.eig_a.p, n, w.eig_w.p, w.eig_work.p, w.eig_info.p));
  rocblas_int info = 0;
  HIPCHK(hipMemcpyAsync(&info, w.eig_info.p, sizeof(info), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (info != 0) return fail(PWH_ERR_LAPACK, "zheevd info " + std::to_string(info));
  HIPCHK(hipMemcpy2DAsync(A, sizeof(cx<double>) * ldA, w.eig_a.p, sizeof(cx<double>) * n,
                          sizeof(cx<double>) * n, n, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(lambda, w.eig_w.p, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

// The small solver on device memory with its scratch in the library's device-resident
// array: no allocation, no synchronisation, no rocSOLVER fallback (info reports a
// non-finite value instead).  Nothing here may allocate device memory, so neither the
// library's init nor its scratch buffers are touched.
int pwh_eig_zheevd_noalloc(double* A_device, int ldA, double* lambda_device, int n, int* info_device,
                           void* stream) {
  if (!A_device || !lambda_device || !info_device || n < 1 || ldA < n)
    return fail(PWH_ERR_ARG, "bad eig_zheevd_noalloc arguments");
  if (n > kHeevSmallMax)
    return fail(PWH_ERR_ARG, "eig_zheevd_noalloc: n = " + std::to_string(n) + " exceeds " +
                                 std::to_string(kHeevSmallMax) + ", the largest matrix its device scratch holds");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g.stream;
  if (!st) return fail(PWH_ERR_ARG, "eig_zheevd_noalloc: no stream given and pwh_init not called");
  double* sc = nullptr;
  HIPCHK(heev_small_static_scratch(&sc));
  HIPCHK(launch_heev_small(reinterpret_cast<cx<double>*>(A_device), ldA, n, lambda_device, info_device, sc, st));
  return PWH_OK;
}

// Density of nb states (EvalRhoPsi.cpp:83-157) held in sphere order in Xw, which the
// call modifies (the last plane wave is zeroed).  stage: the states in reference order to
// permute into Xw first, or null when Xw already holds them.  The density is added to the
// host array rhoe, or, with rho_dev, written to that device vector (overwritten).
static int eval_rho_core(pwh_system* s, const cx<double>* stage, cx<double>* Xw, int nb, const double* w,
                         double omega, double* rhoe, double* ekin, double* charge, double* rho_dev = nullptr) {
  const int ngw = s->ngw, ldp = s->ngw_pad;
  const size_t nnr = (size_t)s->nnr;
  CHK(s->rho_q.ensure((size_t)nb * kBandSlices));
  CHK(s->rho_k.ensure((size_t)nb * kBandSlices));
  CHK(s->W64.ensure((size_t)nb * s->nr[2] * s->maps.ncol_pad));
  std::vector<double> wo(nb);
  for (int b = 0; b < nb; ++b) wo[b] = w[b] / omega;
  CHK(upload(s->rho_w, wo.data(), wo.size()));
  const int groups = std::max(1, std::min(nb, 512 / s->nr[2]));
  CHK(s->rho_part.ensure((size_t)groups * nnr));
  CHK(s->rho.ensure(nnr));
  int chunk;
  CHK(split_prepare<double>(s, nb, &chunk));
  int last = 0;
  for (int k = 0; k < ngw; ++k)
    if (s->perm[k] == ngw - 1) last = k;
  {
    Timed tall(TM_RHO, g.stream);
    {
      Timed t(TM_RHO_SUMS, g.stream);
      if (stage) HIPCHK(permute_in_f64(stage, ngw, s->d_perm.p, ngw, ldp, Xw, nb, g.stream));
      HIPCHK(hipMemset2DAsync(Xw + last, sizeof(cx<double>) * ldp, 0, sizeof(cx<double>), nb, g.stream));
      HIPCHK(band_sums(Xw, ldp, ngw, s->kin64.p, nb, s->rho_q.p, s->rho_k.p, g.stream));
    }
    {
      // inverse transform along i3, then the planes with w |psi|^2 summed over bands (FP64)
      Timed t(TM_RHO_ZINV, g.stream);
      HIPCHK(launch_zinv<double>(s->maps, s->tw64[2].p, Xw, ldp, s->W64.p, nb, g.stream));
    }
    Timed t(TM_RHO_PLANE, g.stream);
    HIPCHK(launch_plane_rho(s->maps, s->tw64[0].p, s->tw64[1].p, s->W64.p, nb, groups, s->rho_w.p,
                            s->rho_part.p, s->rho.p, s->Ysp64.p, chunk, g.stream));
  }
  const size_t nsl = (size_t)nb * kBandSlices;
  std::vector<double> rho(rho_dev ? 0 : nnr), qs(nsl), ks(nsl);
  if (rho_dev)
    HIPCHK(hipMemcpyAsync(rho_dev, s->rho.p, sizeof(double) * nnr, hipMemcpyDeviceToDevice, g.stream));
  else
    HIPCHK(hipMemcpyAsync(rho.data(), s->rho.p, sizeof(double) * nnr, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(qs.data(), s->rho_q.p, sizeof(double) * nsl, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(ks.data(), s->rho_k.p, sizeof(double) * nsl, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (!rho_dev)
    for (size_t i = 0; i < nnr; ++i) rhoe[i] += rho[i];
  // sum_b w_b/2 sum_G XKG |c|^2 TPIBA2 = sum_b w_b sum_G kin |c|^2 (kin = XKG TPIBA2 / 2)
  double e = 0.0;
  for (int b = 0; b < nb; ++b) {
    double q = 0.0, k = 0.0;
    for (int p = 0; p < kBandSlices; ++p) {
      q += qs[(size_t)b * kBandSlices + p];
      k += ks[(size_t)b * kBandSlices + p];
    }
    charge[b] = q;
    e += w[b] * k;
  }
  *ekin = e;
  return PWH_OK;
}

int pwh_eval_rho_psi(pwh_system* s, const double* eigvec, int ld, int nstates, const double* w, double omega,
                     double* rhoe, double* ekin, double* charge) {
  CHK(check_ld(s, nstates));
  if (!eigvec || !w || !rhoe || !ekin || !charge || ld < s->ngw || !(omega > 0.0))
    return fail(PWH_ERR_ARG, "bad eval_rho_psi arguments");
  const int nb = nstates, ngw = s->ngw;
  // coefficients in sphere order; the reference's loop stops at ig < NGW_ik, so the last
  // plane wave contributes to neither the density nor the sums (EvalRhoPsi.cpp:117)
  CHK(s->stage.ensure((size_t)ngw * nb));
  HIPCHK(hipMemcpy2DAsync(s->stage.p, sizeof(cx<double>) * ngw, eigvec, sizeof(cx<double>) * ld,
                          sizeof(cx<double>) * ngw, nb, hipMemcpyHostToDevice, g.stream));
  CHK(s->X.ensure((size_t)nb * s->ngw_pad));
  return eval_rho_core(s, s->stage.p, s->X.p, nb, w, omega, rhoe, ekin, charge);
}

int pwh_resident_eval_rho_psi(pwh_system* s, int nstates, const double* w, double omega, double* rhoe,
                              double* ekin, double* charge) {
  CHK(check_ld(s, nstates));
  if (!w || !rhoe || !ekin || !charge || !(omega > 0.0)) return fail(PWH_ERR_ARG, "bad eval_rho_psi arguments");
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  // a copy of the first nstates rows: the resident block keeps its last plane wave
  const size_t n = (size_t)nstates * s->ngw_pad;
  CHK(s->dtmp.ensure(n));
  HIPCHK(hipMemcpyAsync(s->dtmp.p, s->X.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  return eval_rho_core(s, nullptr, s->dtmp.p, nstates, w, omega, rhoe, ekin, charge);
}

int pwh_lanczos(pwh_system* s, double* v0, int k, double* T, int ldT, double* f) {
  CHK(check_ld(s, 1));
  if (!v0 || !T || !f || k < 1 || ldT < k) return fail(PWH_ERR_ARG, "bad lanczos arguments");
  CHK(load_block(s, v0, 1));
  CHK(s->HX.ensure((size_t)s->ngw_pad));
  double fn2 = 0;
  CHK(lanczos_dev(s, s->X.p, k, T, ldT, s->HX.p, &fn2));
  CHK(store_block(s, s->HX.p, f, 1));
  return store_block(s, s->X.p, v0, 1);
}

int pwh_rayleigh_ritz(pwh_system* s, double* X, int nstates, double* eig) {
  CHK(check_ld(s, nstates));
  if (!X || !eig) return fail(PWH_ERR_ARG, "null argument");
  CHK(load_block(s, X, nstates));
  CHK(rayleigh_ritz_dev(s, s->X.p, nstates));
  HIPCHK(hipMemcpyAsync(eig, s->evals.p, sizeof(double) * nstates, hipMemcpyDeviceToHost, g.stream));
  return store_block(s, s->X.p, X, nstates);
}

int pwh_residual_norms(pwh_system* s, const double* X, int nstates, double* res) {
  CHK(check_ld(s, nstates));
  if (!X || !res) return fail(PWH_ERR_ARG, "null argument");
  CHK(load_block(s, X, nstates));
  return residuals_dev(s, s->X.p, nstates, res);
}

// X: host block to load and store, or null for the resident block (first nstates rows)
static int chfsi_impl(pwh_system* s, double* X, int nstates, double* v0, pwh_chfsi_opts* o, double* eig,
                      double* res) {
  CHK(check_ld(s, nstates));
  if (!o || !eig || !res) return fail(PWH_ERR_ARG, "null argument");
  if (!X && nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  if (o->degree < 1 || o->niter < 0) return fail(PWH_ERR_ARG, "bad degree / niter");
  if (o->precision < PWH_FP64 || o->precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  if (o->lanczos_k > 0 && !v0) return fail(PWH_ERR_ARG, "Lanczos needs v0");
  const int nb = nstates, NX = nstates - 2;
  const double t0 = now_s();
  o->t_filter_s = o->t_ortho_s = o->t_rr_s = o->t_lanczos_s = 0;
  o->band_applies = 0;
  double lb = o->lb, ub = o->ub;
  std::vector<double> T;
  int kref = std::max(2 * NX, 10);
  // Lanczos bounds (ChebySCFLoop.cpp:115,176-196) on a separate vector slot
  Buf<cx<double>> dv0, df;
  auto run_lanczos = [&](int k, int ldT, double* fn2) -> int {
    T.assign((size_t)ldT * k, 0.0);
    CHK(dv0.ensure(s->ngw_pad));
    CHK(df.ensure(s->ngw_pad));
    CHK(upload(s->stage, reinterpret_cast<const cx<double>*>(v0), s->ngw));
    HIPCHK(permute_in_f64(s->stage.p, s->ngw, s->d_perm.p, s->ngw, s->ngw_pad, dv0.p, 1, g.stream));
    const double ta = now_s();
    CHK(lanczos_dev(s, dv0.p, k, T.data(), ldT, df.p, fn2));
    o->t_lanczos_s += now_s() - ta;
    o->band_applies += k;
    return PWH_OK;
  };
  int rc = PWH_OK;
  // the reference bound rule (bounds_mode 1) runs a Lanczos after every RR, whose dot
  // products use s->S, where the eigenvector half writes: no overlap then
  const bool overlap = chfsi_overlap_on(nb) && o->bounds_mode != 1;
  bool pending = false;
  if (o->lanczos_k > 0) {
    const int k = (o->bounds_mode == 1) ? kref : o->lanczos_k;
    double fn2 = 0;
    if ((rc = run_lanczos(k, k, &fn2))) goto done;
    std::vector<double> w;
    if ((rc = sym_eigvals_upper(T.data(), k, k, w))) goto done;
    if (o->bounds_mode == 1) {
      if (k < 2 * NX) { rc = fail(PWH_ERR_ARG, "reference bounds need k >= 2 NX"); goto done; }
      lb = w[NX + 1];
      ub = w[2 * NX - 1];
    } else {
      ub = w[k - 1] + std::sqrt(fn2);
      const double frac = std::min(0.5, std::max(0.02, std::pow((double)nb / s->ngw, 2.0 / 3.0)));
      lb = w[0] + frac * (ub - w[0]);
    }
  }
  if (X && (rc = load_block(s, X, nb))) goto done;
  for (int it = 0; it < o->niter; ++it) {
    double ta = now_s();
    // from the second iteration on, H X of the block is that of the last RR: rotated, or,
    // while that RR's eigenvectors are still being computed on the third stream, the
    // unrotated pair, whose filtered block is rotated once they are ready
    if ((rc = filter_prec(s, s->X.p, nb, o->degree, lb, ub, o->precision, it > 0 ? s->HX.p : nullptr)))
      goto done;
    if (pending) {
      pending = false;
      if ((rc = rr_finish(s, s->X.p, nb, false, eig))) goto done;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    o->t_filter_s += now_s() - ta;
    o->band_applies += (long long)o->degree * nb;
    ta = now_s();
    if ((rc = sys_cholqr2(s, s->X.p, nb, true))) goto done;
    HIPCHK(hipStreamSynchronize(g.stream));
    o->t_ortho_s += now_s() - ta;
    ta = now_s();
    if (overlap) {
      if ((rc = rr_begin(s, s->X.p, nb, eig))) goto done;
      if (it + 1 < o->niter) pending = true;
      else if ((rc = rr_finish(s, s->X.p, nb, true, eig))) goto done;
    } else {
      if ((rc = rayleigh_ritz_dev(s, s->X.p, nb))) goto done;
      HIPCHK(hipMemcpyAsync(eig, s->evals.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g.stream));
      HIPCHK(hipStreamSynchronize(g.stream));
    }
    o->t_rr_s += now_s() - ta;
    o->band_applies += nb;
    // bounds update (ChebySCFLoop.cpp:355-368)
    lb = eig[nb - 1];
    if (o->bounds_mode == 1) {
      const int kk = (kref > 10) ? 10 : kref;
      double fn2 = 0;
      if ((rc = run_lanczos(kk, kref, &fn2))) goto done;
      ub = induced_norm_sq(T.data(), kref, kk) + std::sqrt(fn2);
    }
  }
  // residual report of the final block
  if ((rc = residuals_dev(s, s->X.p, nb, res, o->niter > 0))) goto done;
  if (X) {
    if ((rc = store_block(s, s->X.p, X, nb))) goto done;
  } else {
    // keep the resident FP32 copy (pwh_resident_apply) equal to the new block
    const size_t n = (size_t)s->resident_nb * s->ngw_pad;
    if ((rc = s->X32.ensure(n))) goto done;
    HIPCHK((convert_cx<double, float>(s->X.p, s->X32.p, n, g.stream)));
  }
  o->lb = lb;
  o->ub = ub;
  o->t_total_s = now_s() - t0;
done:
  // an error left an eigenvector half running: let it finish before the buffers are reused
  if (pending && s->eig_stream) hipStreamSynchronize(s->eig_stream);
  dv0.release();
  df.release();
  return rc;
}

int pwh_chfsi(pwh_system* s, double* X, int nstates, double* v0, pwh_chfsi_opts* o, double* eig,
              double* res) {
  if (!X) return fail(PWH_ERR_ARG, "null argument");
  return chfsi_impl(s, X, nstates, v0, o, eig, res);
}

int pwh_resident_chfsi(pwh_system* s, int nstates, double* v0, pwh_chfsi_opts* o, double* eig, double* res) {
  return chfsi_impl(s, nullptr, nstates, v0, o, eig, res);
}

// ---------------------------------------------------------------------------
// Block eigensolvers of the SCF driver on the device: davidson_cpu.cpp:9-247 and
// lobpcg_cpu.cpp:32-389.  Blocks stay in HBM in sphere order for the whole solve; the
// host sees the per-band residual norms once per iteration (the convergence tests).
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

constexpr double kDiagTol = 1.0e-6;                     // davidson_cpu.cpp:85, lobpcg_cpu.cpp:35
constexpr double kZhegvSmall = 12.220446049250313e-16;  // eig_zhegv.cpp:8

struct DiagRun {
  pwh_system* s;
  pwh_diag_opts* o;
  int ld, ngw;
  size_t pe;   // partial-sum elements of the largest projection
  void report(int ev, int a, int b, double v) const {
    if (o->report) o->report(o->report_ctx, ev, a, b, v);
  }
  // C (column-major, ldc) = A^H B for blocks of ma and nb bands
  int cn(const cx<double>* A, int ma, const cx<double>* B, int nb, cx<double>* C, int ldc) const {
    HIPCHK(gemm_cn_f64(A, ld, ma, B, ld, nb, ngw, C, ldc, s->part64.p, pe, g.stream));
    return PWH_OK;
  }
  // out (nb bands) = A (ja bands) M[:, 0:nb] with M column-major (ldm)
  int nn(const cx<double>* A, int ja, const cx<double>* M, int ldm, int nb, cx<double>* out) const {
    HIPCHK(gemm_nn_f64(A, ld, ja, M, ldm, nb, ngw, out, ld, g.stream));
    return PWH_OK;
  }
  int copy(cx<double>* dst, const cx<double>* src, int nb) const {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(cx<double>) * nb * ld, hipMemcpyDeviceToDevice, g.stream));
    return PWH_OK;
  }
  int info_value(int* v) const {
    rocblas_int h = 0;
    HIPCHK(hipMemcpyAsync(&h, s->info.p, sizeof(h), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *v = h;
    return PWH_OK;
  }
};

// C = op(A) B on column-major N x N device matrices (rocBLAS, host scalars)
int small_zgemm(rocblas_operation ta, int N, const cx<double>* A, const cx<double>* B, cx<double>* C) {
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  RBCHK(rocblas_zgemm(g.rb, ta, rocblas_operation_none, N, N, N, &one,
                      reinterpret_cast<const rocblas_double_complex*>(A), N,
                      reinterpret_cast<const rocblas_double_complex*>(B), N, &zero,
                      reinterpret_cast<rocblas_double_complex*>(C), N));
  return PWH_OK;
}

// eig_zhegv.cpp:5-149 on column-major N x N device matrices: B = V diag(b) V^H (upper
// triangle), columns with b > SMALL scaled by 1/sqrt(b), A' = V^H A V (upper triangle),
// A' = Z diag(e) Z^H, E = V Z.  A and B are overwritten.  Overlap eigenvalues <= SMALL are
// reported (the reference prints them); a failed eigensolve is reported and leaves E as
// it was, as the reference returns early.
int zhegv_dev(const DiagRun& r, cx<double>* A, cx<double>* B, int N, cx<double>* E, double* e) {
  pwh_system* s = r.s;
  CHK(s->dbo.ensure(N));
  CHK(s->rwork.ensure(N));
  CHK(s->dT1.ensure((size_t)N * N));
  HIPCHK(fill_lower(B, N, N, g.stream));
  CHK(hermitian_eig_impl(B, N, s->dbo.p, s->rwork.p, s->info.p));
  int info = 0;
  CHK(r.info_value(&info));
  if (info != 0) {
    r.report(PWH_DIAG_EIG_FAILED, info, N, 0.0);
    return PWH_OK;
  }
  std::vector<double> b(N);
  HIPCHK(hipMemcpyAsync(b.data(), s->dbo.p, sizeof(double) * N, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  int nn = 0;
  for (int i = 0; i < N; ++i) {
    if (b[i] > kZhegvSmall) ++nn;
    else r.report(PWH_DIAG_SMALL_OVERLAP, i + 1, N, b[i]);
  }
  if (nn < N) r.report(PWH_DIAG_RANK, nn, N, 0.0);
  HIPCHK(zhegv_scale(B, N, N, s->dbo.p, g.stream));
  CHK(small_zgemm(rocblas_operation_none, N, A, B, s->dT1.p));
  CHK(small_zgemm(rocblas_operation_conjugate_transpose, N, B, s->dT1.p, A));
  HIPCHK(fill_lower(A, N, N, g.stream));
  CHK(hermitian_eig_impl(A, N, e, s->rwork.p, s->info.p));
  CHK(r.info_value(&info));
  if (info != 0) {
    r.report(PWH_DIAG_EIG_FAILED, info, N, 0.0);
    return PWH_OK;
  }
  return small_zgemm(rocblas_operation_none, N, B, A, E);
}

// A <- A C^{-1}, HA <- HA C^{-1} with C^H C = sym(A^H A) (zpotrf 'U' + ztrsm 'R','U','N','N',
// lobpcg_cpu.cpp:138-166, 330-366); the inverse factor multiplies on the MFMA GEMM
int chol_right(const DiagRun& r, cx<double>* A, cx<double>* HA, int n) {
  pwh_system* s = r.s;
  cx<double>* C = s->S.p;
  CHK(r.cn(A, n, A, n, C, n));
  HIPCHK(symmetrize(C, n, n, g.stream));
  if (n <= kCholSmallMax) {
    HIPCHK(launch_chol_inv_small(C, n, n, reinterpret_cast<int*>(s->info.p), g.stream));
  } else {
    RBCHK(rocsolver_zpotrf(g.rb, rocblas_fill_upper, n, reinterpret_cast<rocblas_double_complex*>(C), n,
                           s->info.p));
  }
  int info = 0;
  CHK(r.info_value(&info));
  if (info != 0)
    return fail(PWH_ERR_LAPACK, "Cholesky factorisation of the block overlap failed (zpotrf info " +
                                    std::to_string(info) + ")");
  if (n > kCholSmallMax) {
    RBCHK(rocsolver_ztrtri(g.rb, rocblas_fill_upper, rocblas_diagonal_non_unit, n,
                           reinterpret_cast<rocblas_double_complex*>(C), n, s->info.p));
    HIPCHK(zero_lower(C, n, n, g.stream));
  }
  CHK(r.nn(A, n, C, n, n, s->dtmp.p));
  CHK(r.copy(A, s->dtmp.p, n));
  CHK(r.nn(HA, n, C, n, n, s->dtmp.p));
  return r.copy(HA, s->dtmp.p, n);
}

// Shared set-up: buffers for nq blocks of n bands, X (host, reference order) into the
// first block, prec into sphere order.
int diag_prepare(pwh_system* s, const double* X, int n, int nq, const double* prec, DiagRun* r) {
  const int ld = s->ngw_pad;
  const size_t m = (size_t)nq * n;
  CHK(s->dq.ensure(m * ld));
  CHK(s->dhq.ensure(m * ld));
  CHK(s->dtmp.ensure((size_t)n * ld));
  CHK(s->dT.ensure(m * m));
  CHK(s->dG.ensure(m * m));
  CHK(s->dE.ensure(m * m));
  CHK(s->dt1.ensure((size_t)n * n));
  CHK(s->S.ensure((size_t)n * n));
  CHK(s->dlam.ensure(m + n));   // n estimates + the m eigenvalues of the projected problem
  CHK(s->dnrm.ensure(n));
  CHK(s->dpart.ensure((size_t)n * kBandSlices));
  CHK(s->info.ensure(2));
  r->pe = std::max(proj_part_elems((int)m, (int)m, s->ngw), proj_part_elems(n, n, s->ngw));
  CHK(s->part64.ensure(r->pe));
  CHK(s->A64.ensure((size_t)n * ld));
  std::vector<double> p(ld, 0.0);
  for (int k = 0; k < s->ngw; ++k) p[k] = prec[s->perm[k]];
  CHK(upload(s->dprec, p.data(), p.size()));
  CHK(upload(s->stage, reinterpret_cast<const cx<double>*>(X), (size_t)s->ngw * n));
  HIPCHK(permute_in_f64(s->stage.p, s->ngw, s->d_perm.p, s->ngw, ld, s->dq.p, n, g.stream));
  r->s = s;
  r->ld = ld;
  r->ngw = s->ngw;
  return PWH_OK;
}

int download_vec(double* h, const double* d, int n) {
  HIPCHK(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int check_diag_args(pwh_system* s, double* X, int n, const double* prec, pwh_diag_opts* o, double* ev,
                    double* res) {
  CHK(check_ld(s, n));
  if (!X || !prec || !o || !ev || !res) return fail(PWH_ERR_ARG, "null argument");
  if (o->max_iter < 1) return fail(PWH_ERR_ARG, "max_iter < 1");
  return PWH_OK;
}

}  // namespace

extern "C" {

int pwh_davidson(pwh_system* s, double* X, int nstates, const double* prec, pwh_diag_opts* o, double* evals,
                 double* res) {
  CHK(check_diag_args(s, X, nstates, prec, o, evals, res));
  const double t0 = now_s();
  const int n = nstates, n2 = 2 * nstates;
  DiagRun r{};
  r.o = o;
  CHK(diag_prepare(s, X, n, 2, prec, &r));
  const int ld = r.ld;
  cx<double>*Q = s->dq.p, *Xd = Q, *R = Q + (size_t)n * ld;
  cx<double>*HQ = s->dhq.p, *HX = HQ, *HR = HQ + (size_t)n * ld;
  cx<double>*h = s->dT.p, *ov = s->dG.p, *E = s->dE.p;
  double *ev = s->dlam.p, *dr = s->dnrm.p, *part = s->dpart.p;
  o->band_applies = 0;
  o->iterations = 0;
  o->reason = 0;
  // orthonormal start block, H X, Rayleigh quotients and residuals (davidson_cpu.cpp:47-71)
  CHK(cholqr2(Xd, ld, s->ngw, n, s->A64.p, s->S.p, s->part64.p, r.pe, s->info.p));
  CHK(hpsi<double>(s, Xd, HX, n, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  o->band_applies += n;
  HIPCHK(dotc_re(Xd, HX, ld, s->ngw, n, part, ev, g.stream));
  HIPCHK(resid_block(Xd, HX, ev, ld, s->ngw, n, 0, R, part, dr, g.stream));
  CHK(download_vec(res, dr, n));
  int nconv = 0;
  double rnorm = 1.0;
  for (int istep = 1; istep <= o->max_iter; ++istep) {
    o->iterations = istep;
    r.report(PWH_DIAG_ITER, istep, nconv, rnorm);
    // residuals normalised and preconditioned, H R (:91-123)
    HIPCHK(scale_prec(R, ld, s->ngw, n, dr, s->dprec.p, g.stream));
    CHK(hpsi<double>(s, R, HR, n, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
    o->band_applies += n;
    // reduced Hamiltonian and overlap, 2n x 2n (:125-184)
    if (istep == 1) CHK(r.cn(Xd, n, HX, n, h, n2));
    else HIPCHK(set_diag_block(h, n2, n, ev, g.stream));
    CHK(r.cn(Q, n2, HR, n, h + (size_t)n * n2, n2));     // <x|H|r>, <r|H|r>
    CHK(r.cn(HR, n, Xd, n, h + n, n2));                  // (H r)^H x
    HIPCHK(set_diag_block(ov, n2, n, nullptr, g.stream));
    CHK(r.cn(Q, n2, R, n, ov + (size_t)n * n2, n2));     // <x|r>, <r|r>
    CHK(r.cn(R, n, Xd, n, ov + n, n2));                  // <r|x>
    CHK(zhegv_dev(r, h, ov, n2, E, ev));
    // X <- [X R] E(:, 1:n), HX <- [HX HR] E(:, 1:n) (:189-213)
    CHK(r.nn(Q, n2, E, n2, n, s->dtmp.p));
    CHK(r.copy(Xd, s->dtmp.p, n));
    CHK(r.nn(HQ, n2, E, n2, n, s->dtmp.p));
    CHK(r.copy(HX, s->dtmp.p, n));
    HIPCHK(resid_block(Xd, HX, ev, ld, s->ngw, n, 0, R, part, dr, g.stream));
    CHK(download_vec(res, dr, n));
    // convergence tests (:215-236)
    nconv = 0;
    for (int i = 0; i < n; ++i) nconv += res[i] < kDiagTol;
    if (nconv >= n - o->nempty) { o->reason = 1; break; }
    rnorm = 0.0;
    for (int i = 0; i < n; ++i) rnorm += res[i];
    rnorm /= n;
    if (rnorm < kDiagTol / 10.0) { o->reason = 2; break; }
  }
  o->nconv = nconv;
  CHK(download_vec(evals, ev, n));
  CHK(store_block(s, Xd, X, n));
  o->t_total_s = now_s() - t0;
  return PWH_OK;
}

int pwh_lobpcg(pwh_system* s, double* X, int nstates, const double* prec, pwh_diag_opts* o, double* lambda,
               double* res) {
  CHK(check_diag_args(s, X, nstates, prec, o, lambda, res));
  const double t0 = now_s();
  const int n = nstates, n2 = 2 * nstates, n3 = 3 * nstates;
  DiagRun r{};
  r.o = o;
  CHK(diag_prepare(s, X, n, 3, prec, &r));
  const int ld = r.ld;
  cx<double>*Q = s->dq.p, *Xd = Q, *W = Q + (size_t)n * ld, *P = Q + (size_t)n2 * ld;
  cx<double>*HQ = s->dhq.p, *HX = HQ, *HW = HQ + (size_t)n * ld, *HP = HQ + (size_t)n2 * ld;
  cx<double>*T = s->dT.p, *G = s->dG.p, *U = s->dE.p, *t1 = s->dt1.p;
  double *lam = s->dlam.p, *dr = s->dnrm.p, *part = s->dpart.p;
  o->band_applies = 0;
  o->iterations = 0;
  o->reason = 0;
  // orthonormal start block and H X (lobpcg_cpu.cpp:80-91)
  CHK(cholqr2(Xd, ld, s->ngw, n, s->A64.p, s->S.p, s->part64.p, r.pe, s->info.p));
  CHK(hpsi<double>(s, Xd, HX, n, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
  o->band_applies += n;
  int nconv = 0;
  double rnorm = 0.0;
  const int iters = std::max(1, o->max_iter);
  for (int it = 1; it <= iters; ++it) {
    o->iterations = it;
    if (it > 1) r.report(PWH_DIAG_ITER, it, nconv, rnorm);
    // X^H H X, W = HX - X (X^H H X), its eigenvalues, residual norms (:93-124, 189-222)
    CHK(r.cn(Xd, n, HX, n, t1, n));
    CHK(r.nn(Xd, n, t1, n, n, s->dtmp.p));
    HIPCHK(resid_block(s->dtmp.p, HX, nullptr, ld, s->ngw, n, 1, W, part, dr, g.stream));
    HIPCHK(fill_lower(t1, n, n, g.stream));
    CHK(s->rwork.ensure(n));
    CHK(hermitian_eig_impl(t1, n, lam, s->rwork.p, s->info.p));
    int info = 0;
    CHK(r.info_value(&info));
    if (info != 0) r.report(PWH_DIAG_EIG_FAILED, info, n, 0.0);
    CHK(download_vec(res, dr, n));
    CHK(download_vec(lambda, lam, n));
    nconv = 0;
    rnorm = 0.0;
    for (int i = 0; i < n; ++i) {
      nconv += res[i] < kDiagTol;
      rnorm += res[i];
    }
    rnorm /= n;
    if (nconv >= n - o->nempty) { o->reason = 1; break; }
    if (rnorm < kDiagTol) { o->reason = 2; break; }
    // preconditioned W, H W, W orthonormalised by its Cholesky factor (:126-166)
    HIPCHK(scale_prec(W, ld, s->ngw, n, nullptr, s->dprec.p, g.stream));
    CHK(hpsi<double>(s, W, HW, n, 1.0, 0.0, 0.0, nullptr, PREC_FP64));
    o->band_applies += n;
    CHK(chol_right(r, W, HW, n));
    // Rayleigh-Ritz on [X W] (first iteration) or [X W P] (:168-186, 268-287)
    const int m = it == 1 ? n2 : n3;
    CHK(r.cn(Q, m, HQ, m, T, m));
    HIPCHK(symmetrize(T, m, m, g.stream));
    CHK(r.cn(Q, m, Q, m, G, m));
    HIPCHK(symmetrize(G, m, m, g.stream));
    CHK(s->dbo.ensure(m));
    CHK(zhegv_dev(r, T, G, m, U, s->dlam.p + n));
    // X <- Q U(:, 1:n), HX <- HQ U(:, 1:n); P from W and the previous P (:188-201, 289-328)
    CHK(r.nn(Q, m, U, m, n, s->dtmp.p));
    CHK(r.copy(Xd, s->dtmp.p, n));
    CHK(r.nn(HQ, m, U, m, n, s->dtmp.p));
    CHK(r.copy(HX, s->dtmp.p, n));
    if (it == 1) {
      CHK(r.copy(P, W, n));
      CHK(r.copy(HP, HW, n));
    } else {
      CHK(r.nn(W, n2, U + n, m, n, s->dtmp.p));
      CHK(r.copy(P, s->dtmp.p, n));
      CHK(r.nn(HW, n2, U + n, m, n, s->dtmp.p));
      CHK(r.copy(HP, s->dtmp.p, n));
      CHK(chol_right(r, P, HP, n));
    }
  }
  o->nconv = nconv;
  CHK(store_block(s, Xd, X, n));
  o->t_total_s = now_s() - t0;
  return PWH_OK;
}

// ---------------------------------------------------------------------------
// SCF steps around the eigensolver: EvalEnergy.cpp, xc_PZ.cpp, the density mixers,
// NormalizeRho.cpp and EvalENL.cpp, FP64 on the device.  The entry points take host
// arrays as the reference's do; reductions are per-block partials added on the host in
// block order (deterministic).
// ---------------------------------------------------------------------------
}  // extern "C"

struct pwh_grid {
  int nr[3] = {0, 0, 0};
  size_t nnr = 0;
  int ng = 0, ngw = 0;
  double tpiba2 = 0, omega = 0;
  Buf<int> gidx, kidx;
  Buf<double> g2, xkg;
  Buf<cx<double>> vp, rp, tw[3], c;
};

namespace {

// device vectors of the grid-sized steps (one set for the library; the library runs one
// caller thread)
struct ScfWork {
  Buf<double> v[24];
  Buf<double> part, vxc, eps;
  Buf<unsigned char> neg;
  Buf<double> small;
  Buf<rocblas_int> ipiv, info;
};
ScfWork& scfw() {
  static ScfWork* w = new ScfWork();
  return *w;
}

double sum_partials(const double* d, int n, double* host) {
  hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream);
  hipStreamSynchronize(g.stream);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += host[i];
  return s;
}

// sum over the grid of x*y (device vectors), block partials added in order
int dot_dev(const double* x, const double* y, size_t n, double* out) {
  ScfWork& w = scfw();
  CHK(w.part.ensure(4 * kScfBlocks));
  HIPCHK(dot_partials(x, y, n, w.part.p, g.stream));
  std::vector<double> h(kScfBlocks);
  *out = sum_partials(w.part.p, kScfBlocks, h.data());
  return PWH_OK;
}

// out = sum_k w_k v_k; with p: returns sum (p - q)^2 in *sq
int comb_dev(std::initializer_list<std::pair<double, const double*>> terms, double* out, size_t n,
             double* out2 = nullptr, const double* p = nullptr, const double* q = nullptr, double* sq = nullptr) {
  Comb cb{};
  cb.nt = 0;
  for (auto& t : terms) {
    if (cb.nt == kMaxTerms) return fail(PWH_ERR_ARG, "too many terms");
    cb.w[cb.nt] = t.first;
    cb.v[cb.nt] = t.second;
    ++cb.nt;
  }
  cb.out = out;
  cb.out2 = out2;
  cb.p = p;
  cb.q = q;
  ScfWork& w = scfw();
  CHK(w.part.ensure(4 * kScfBlocks));
  HIPCHK(comb(cb, n, p ? w.part.p : nullptr, g.stream));
  if (p && sq) {
    std::vector<double> h(kScfBlocks);
    *sq = sum_partials(w.part.p, kScfBlocks, h.data());
  }
  return PWH_OK;
}

int scf_up(Buf<double>& b, const double* h, size_t n) { return upload(b, h, n); }
int scf_down(double* h, const double* d, size_t n) {
  HIPCHK(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream));
  return PWH_OK;
}
int scf_sync() {
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

// 3D FFT of the grid's complex work array (fft_fftw3.cpp conventions)
int grid_fft(pwh_grid* gr, cx<double>* c, bool inverse) {
  const int sign = inverse ? +1 : -1;
  const double scale = inverse ? 1.0 : 1.0 / (double)gr->nnr;
  const int n1 = gr->nr[0], n2 = gr->nr[1], n3 = gr->nr[2];
  HIPCHK(launch_fft_lines(c, n3, n1 * n2, n1 * n2, gr->tw[2].p, sign, 1.0, g.stream));
  HIPCHK(launch_fft_lines(c, n2, n1, n1 * n3, gr->tw[1].p, sign, 1.0, g.stream));
  HIPCHK(launch_fft_lines(c, n1, 1, n2 * n3, gr->tw[0].p, sign, scale, g.stream));
  return PWH_OK;
}

// KerkerRhoMix.cpp:66-79 on the device vector x (in place)
int kerker_dev(pwh_grid* gr, double* x, double beta, double q0) {
  CHK(gr->c.ensure(gr->nnr));
  HIPCHK(real_to_cx(x, nullptr, gr->nnr, gr->c.p, g.stream));
  CHK(grid_fft(gr, gr->c.p, false));
  HIPCHK(kerker_filter(gr->c.p, gr->kidx.p, gr->xkg.p, gr->ngw, beta, q0, g.stream));
  CHK(grid_fft(gr, gr->c.p, true));
  HIPCHK(cx_real(gr->c.p, gr->nnr, x, g.stream));
  return PWH_OK;
}

// NormalizeRho.cpp:7-33 on a device vector; *diff = electron-count difference
int normalize_dev(double* x, size_t n, double omega, double nel, double* diff) {
  ScfWork& w = scfw();
  CHK(w.part.ensure(4 * kScfBlocks));
  HIPCHK(clip_sum(x, n, w.part.p, g.stream));
  std::vector<double> h(kScfBlocks);
  const double sum = sum_partials(w.part.p, kScfBlocks, h.data());
  const double truesum = (double)n * nel / omega;
  if (sum < 1.0e-15) return fail(PWH_ERR_ARG, "NormalizeRho: sum(rho) < 1e-15");
  if (diff) *diff = (truesum - sum) * omega / (double)n;
  HIPCHK(scale_vec(x, n, truesum, sum, g.stream));   // rho * truesum / sum (NormalizeRho.cpp:29-31)
  return PWH_OK;
}

// Solve A x = b (m x m, column-major) on the device with LU; b is replaced by x
int lu_solve_dev(std::vector<double>& A, std::vector<double>& b, int m) {
  ScfWork& w = scfw();
  CHK(w.small.ensure((size_t)m * m + m));
  CHK(w.ipiv.ensure(m));
  CHK(w.info.ensure(1));
  HIPCHK(hipMemcpyAsync(w.small.p, A.data(), sizeof(double) * m * m, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(w.small.p + (size_t)m * m, b.data(), sizeof(double) * m, hipMemcpyHostToDevice,
                        g.stream));
  RBCHK(rocsolver_dgetrf(g.rb, m, m, w.small.p, m, w.ipiv.p, w.info.p));
  rocblas_int info = 0;
  HIPCHK(hipMemcpyAsync(&info, w.info.p, sizeof(info), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (info != 0) return fail(PWH_ERR_LAPACK, "dgetrf info " + std::to_string(info));
  RBCHK(rocsolver_dgetrs(g.rb, rocblas_operation_none, m, 1, w.small.p, m, w.ipiv.p, w.small.p + (size_t)m * m,
                         m));
  HIPCHK(hipMemcpyAsync(b.data(), w.small.p + (size_t)m * m, sizeof(double) * m, hipMemcpyDeviceToHost,
                        g.stream));
  return scf_sync();
}

// Inverse of the leading m x m block of A (column-major, ld) on the device (dgetrf + dgetri)
int lu_inverse_dev(std::vector<double>& A, int m, int ld) {
  ScfWork& w = scfw();
  CHK(w.small.ensure((size_t)m * m));
  CHK(w.ipiv.ensure(m));
  CHK(w.info.ensure(1));
  std::vector<double> Ap((size_t)m * m);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) Ap[i + (size_t)j * m] = A[i + (size_t)j * ld];
  HIPCHK(hipMemcpyAsync(w.small.p, Ap.data(), sizeof(double) * m * m, hipMemcpyHostToDevice, g.stream));
  RBCHK(rocsolver_dgetrf(g.rb, m, m, w.small.p, m, w.ipiv.p, w.info.p));
  rocblas_int info = 0;
  HIPCHK(hipMemcpyAsync(&info, w.info.p, sizeof(info), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (info != 0) return fail(PWH_ERR_LAPACK, "dgetrf info " + std::to_string(info));
  RBCHK(rocsolver_dgetri(g.rb, m, w.small.p, m, w.ipiv.p, w.info.p));
  HIPCHK(hipMemcpyAsync(&info, w.info.p, sizeof(info), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(Ap.data(), w.small.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (info != 0) return fail(PWH_ERR_LAPACK, "dgetri info " + std::to_string(info));
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) A[i + (size_t)j * ld] = Ap[i + (size_t)j * m];
  return PWH_OK;
}

}  // namespace

extern "C" {

int pwh_grid_create(const pwh_grid_desc* d, pwh_grid** out) {
  if (!d || !out) return fail(PWH_ERR_ARG, "null argument");
  *out = nullptr;
  const int nr[3] = {d->nr1, d->nr2, d->nr3};
  if (!pwh_grid_supported(nr[0], nr[1], nr[2]))
    return fail(PWH_ERR_UNSUPPORTED, "no FFT plan for a length outside 2.." + std::to_string(kFftMaxLen));
  if (d->ng < 1 || !d->n1 || !d->n2 || !d->n3 || !d->g2 || d->nsp < 1 || !d->sfac || !d->vps || !d->rhops ||
      !(d->omega > 0) || !(d->tpiba2 > 0) || d->ngw < 0 || (d->ngw > 0 && (!d->n123 || !d->xkg)))
    return fail(PWH_ERR_ARG, "bad grid description");
  CHK(ensure_init());
  std::unique_ptr<pwh_grid> gr(new pwh_grid());
  for (int k = 0; k < 3; ++k) gr->nr[k] = nr[k];
  gr->nnr = (size_t)nr[0] * nr[1] * nr[2];
  gr->ng = d->ng;
  gr->ngw = d->ngw;
  gr->tpiba2 = d->tpiba2;
  gr->omega = d->omega;
  std::vector<int> gi(d->ng);
  for (int ig = 0; ig < d->ng; ++ig) {
    const int a = d->n1[ig] - 1, b = d->n2[ig] - 1, c = d->n3[ig] - 1;
    if (a < 0 || a >= nr[0] || b < 0 || b >= nr[1] || c < 0 || c >= nr[2])
      return fail(PWH_ERR_ARG, "G-vector grid index out of range");
    gi[ig] = a + nr[0] * (b + nr[1] * c);
  }
  if (gi[0] != 0) return fail(PWH_ERR_ARG, "the first G-vector must be G = 0 (EvalEnergy.cpp:55-57)");
  {
    // each G-vector owns its grid point (a grid too small for the cutoff would alias them)
    std::vector<int> sorted(gi);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return fail(PWH_ERR_ARG, "two G-vectors share a grid point: the grid is too small for the cutoff");
  }
  // species summed in order (EvalEnergy.cpp:51-53, 62-65)
  std::vector<cx<double>> vp(d->ng, mk<double>(0, 0)), rp(d->ng, mk<double>(0, 0));
  for (int ig = 0; ig < d->ng; ++ig)
    for (int is = 0; is < d->nsp; ++is) {
      const size_t k = (size_t)is + (size_t)d->nsp * ig;
      const double sx = d->sfac[2 * k], sy = d->sfac[2 * k + 1];
      vp[ig] = mk<double>(vp[ig].x + sx * d->vps[k], vp[ig].y + sy * d->vps[k]);
      rp[ig] = mk<double>(rp[ig].x + sx * d->rhops[k], rp[ig].y + sy * d->rhops[k]);
    }
  std::vector<int> ki(d->ngw);
  for (int ig = 0; ig < d->ngw; ++ig) {
    if (d->n123[ig] < 1 || d->n123[ig] > gr->nnr) return fail(PWH_ERR_ARG, "N123 out of range");
    ki[ig] = (int)d->n123[ig] - 1;
  }
  CHK(upload(gr->gidx, gi.data(), gi.size()));
  CHK(upload(gr->g2, d->g2, d->ng));
  CHK(upload(gr->vp, vp.data(), vp.size()));
  CHK(upload(gr->rp, rp.data(), rp.size()));
  if (d->ngw > 0) {
    CHK(upload(gr->kidx, ki.data(), ki.size()));
    CHK(upload(gr->xkg, d->xkg, d->ngw));
  }
  for (int k = 0; k < 3; ++k) {
    const auto t = twiddles(nr[k]);
    CHK(upload(gr->tw[k], t.data(), t.size()));
  }
  CHK(gr->c.ensure(gr->nnr));
  HIPCHK(hipStreamSynchronize(g.stream));
  *out = gr.release();
  return PWH_OK;
}

int pwh_grid_destroy(pwh_grid* gr) {
  if (g.stream) hipStreamSynchronize(g.stream);
  delete gr;
  return PWH_OK;
}

int pwh_xc_pz(const double* rho, double* vxc, double* exc, int nnr, int* n_negative) {
  if (!rho || !vxc || !exc || nnr < 1) return fail(PWH_ERR_ARG, "bad xc_pz arguments");
  CHK(ensure_init());
  ScfWork& w = scfw();
  const size_t n = (size_t)nnr;
  CHK(scf_up(w.v[0], rho, n));
  CHK(w.vxc.ensure(n));
  CHK(w.eps.ensure(n));
  CHK(w.neg.ensure(n));
  CHK(w.part.ensure(4 * kScfBlocks));
  HIPCHK(xc_pz(w.v[0].p, n, w.vxc.p, w.eps.p, w.neg.p, nullptr, nullptr, w.part.p, g.stream));
  std::vector<double> h(kScfBlocks);
  const double cnt = sum_partials(w.part.p, kScfBlocks, h.data());
  if (cnt > 0) HIPCHK(xc_pz_fix(w.v[0].p, n, w.neg.p, w.vxc.p, w.eps.p, nullptr, nullptr, g.stream));
  CHK(dot_dev(w.v[0].p, w.eps.p, n, exc));
  CHK(scf_down(vxc, w.vxc.p, n));
  CHK(scf_sync());
  if (n_negative) *n_negative = (int)cnt;
  return PWH_OK;
}

}  // extern "C"

namespace {

// EvalEnergy.cpp:5-163 on the device density rho (read only): V_eff into veff (device), the
// energy terms into e (ekin, enl, esr, eself in)
int eval_energy_core(pwh_grid* gr, const double* rho, double* veff, pwh_energy* e) {
  ScfWork& w = scfw();
  const size_t n = gr->nnr;
  CHK(w.vxc.ensure(n));
  CHK(w.eps.ensure(n));
  CHK(w.neg.ensure(n));
  CHK(w.part.ensure(4 * kScfBlocks));
  // forward transform of the density, G-space sums and the electrostatic potential,
  // inverse transform (EvalEnergy.cpp:38-117)
  std::vector<double> hp(4 * kScfBlocks);
  {
    Timed t(TM_ENERGY, g.stream);
    HIPCHK(real_to_cx(rho, nullptr, n, gr->c.p, g.stream));
    CHK(grid_fft(gr, gr->c.p, false));
    HIPCHK(energy_g(gr->c.p, gr->gidx.p, gr->g2.p, gr->vp.p, gr->rp.p, gr->ng, gr->tpiba2, w.part.p, g.stream));
    HIPCHK(hipMemcpyAsync(hp.data(), w.part.p, sizeof(double) * 4 * kScfBlocks, hipMemcpyDeviceToHost,
                          g.stream));
    CHK(grid_fft(gr, gr->c.p, true));
  }
  CHK(scf_sync());
  double sums[4] = {0, 0, 0, 0};
  for (int q = 0; q < 4; ++q)
    for (int b = 0; b < kScfBlocks; ++b) sums[q] += hp[(size_t)q * kScfBlocks + b];
  // xc_PZ and V_eff = Re(V_es) + V_xc (:124-135)
  {
    Timed t(TM_ENERGY, g.stream);
    HIPCHK(xc_pz(rho, n, w.vxc.p, w.eps.p, w.neg.p, gr->c.p, veff, w.part.p, g.stream));
  }
  std::vector<double> h(kScfBlocks);
  const double cnt = sum_partials(w.part.p, kScfBlocks, h.data());
  if (cnt > 0) HIPCHK(xc_pz_fix(rho, n, w.neg.p, w.vxc.p, w.eps.p, gr->c.p, veff, g.stream));
  double exc = 0;
  CHK(dot_dev(rho, w.eps.p, n, &exc));
  const double om = gr->omega;
  e->exc = exc * om / (double)n;
  e->eht = sums[0] * om + e->esr - e->eself;
  e->epseu = sums[1] * om;
  e->eh_electron = sums[2] * om;
  e->eh_pseudo = sums[3] * om;
  e->etot = e->ekin + e->eht + e->epseu + e->enl + e->exc;
  e->n_negative = (int)cnt;
  return PWH_OK;
}

}  // namespace

extern "C" {

int pwh_eval_energy(pwh_grid* gr, double* rhoe, pwh_energy* e) {
  if (!gr || !rhoe || !e) return fail(PWH_ERR_ARG, "null argument");
  ScfWork& w = scfw();
  const size_t n = gr->nnr;
  CHK(scf_up(w.v[0], rhoe, n));
  CHK(w.v[1].ensure(n));
  CHK(eval_energy_core(gr, w.v[0].p, w.v[1].p, e));
  CHK(scf_down(rhoe, w.v[1].p, n));
  return scf_sync();
}

int pwh_normalize_rho(double* rho, int nnr, double omega, double nel, double* diff) {
  if (!rho || nnr < 1 || !(omega > 0)) return fail(PWH_ERR_ARG, "bad normalize_rho arguments");
  CHK(ensure_init());
  ScfWork& w = scfw();
  CHK(scf_up(w.v[0], rho, nnr));
  CHK(normalize_dev(w.v[0].p, nnr, omega, nel, diff));
  CHK(scf_down(rho, w.v[0].p, nnr));
  return scf_sync();
}

}  // extern "C"

namespace {

// The mixers' device bodies.  R is the reference's rhoe (the new density in, the mixed one
// out); O, F, G, DF, U are the columns of its history arrays (rho_old, f, g, df, u), each a
// device vector of n values.
struct MixCols {
  double* R = nullptr;
  double* O[8] = {};
  double* F[8] = {};
  double* G[2] = {};
  double* DF[8] = {};
  double* U[8] = {};
};

// SimpleRhoMix.cpp:5-20
int mix_simple_core(double beta, const MixCols& c, size_t n, double* d) {
  double sq = 0;
  CHK(comb_dev({{beta, c.R}, {1.0 - beta, c.O[0]}}, c.R, n, c.O[0], c.R, c.O[0], &sq));
  *d = std::sqrt(sq / (double)n);
  return PWH_OK;
}

// KerkerRhoMix.cpp:5-55 (G[0]: a work vector)
int mix_kerker_core(pwh_grid* gr, double beta, double q0, const MixCols& c, double* d) {
  const size_t n = gr->nnr;
  CHK(comb_dev({{1.0, c.R}, {-1.0, c.O[0]}}, c.G[0], n));
  CHK(kerker_dev(gr, c.G[0], beta, q0));
  double sq = 0;
  CHK(comb_dev({{1.0, c.O[0]}, {1.0, c.G[0]}}, c.R, n, c.O[0], c.R, c.O[0], &sq));
  *d = std::sqrt(sq / (double)n);
  return PWH_OK;
}

// AndersonRhoMix.cpp:9-85 (rho_old, f: 3 columns; g: 2)
int mix_anderson_core(int iter, double beta, const MixCols& c, size_t n, double* d) {
  double* R = c.R;
  auto O = [&](int j) { return c.O[j]; };
  auto F = [&](int j) { return c.F[j]; };
  auto G = [&](int j) { return c.G[j]; };
  if (iter == 1) {
    // AndersonRhoMix.cpp:18-23
    CHK(comb_dev({{1.0, R}}, O(0), n));
    CHK(comb_dev({}, F(0), n));
    *d = 1.0;
    return PWH_OK;
  }
  const int j0 = (iter - 1) % 3, j1 = (iter - 2) % 3, j2 = (iter - 3) % 3;
  double sq = 0;
  CHK(comb_dev({{1.0, R}, {-1.0, O(j1)}}, F(j0), n, nullptr, R, O(j1), &sq));
  *d = sq;
  if (iter == 2) {
    CHK(comb_dev({{beta, R}, {1.0 - beta, O(j1)}}, R, n, O(j0)));
    return PWH_OK;
  }
  CHK(comb_dev({{1.0, F(j1)}, {-1.0, F(j0)}}, G(0), n));
  double d11 = 0, r1 = 0;
  CHK(dot_dev(G(0), G(0), n, &d11));
  CHK(dot_dev(F(j0), G(0), n, &r1));
  if (iter == 3) {
    const double t1 = -r1 / d11;
    CHK(comb_dev({{1.0 - t1, O(j1)}, {t1, O(j2)}, {beta, F(j0)}, {beta * t1, G(0)}}, R, n, O(j0)));
    return PWH_OK;
  }
  CHK(comb_dev({{1.0, F(j2)}, {-1.0, F(j0)}}, G(1), n));
  double d22 = 0, d12 = 0, r2 = 0;
  CHK(dot_dev(G(1), G(1), n, &d22));
  CHK(dot_dev(G(0), G(1), n, &d12));
  CHK(dot_dev(F(j0), G(1), n, &r2));
  double t1 = d11 * d22, t2 = t1 - d12 * d12;
  if (std::fabs(t2) > std::fabs(t1) * 1.0e-8) {
    t1 = (-r1 * d22 + r2 * d12) / t2;
    t2 = (r1 * d12 - r2 * d11) / t2;
  } else {
    t1 = 0.0;
    t2 = 0.0;
  }
  // the reference adds t2 g(:,1), not t2 g(:,2) (AndersonRhoMix.cpp:76-78)
  CHK(comb_dev({{1.0 - t1 - t2, O(j1)}, {t1, O(j2)}, {t2, O(j0)}, {beta, F(j0)}, {beta * t1, G(0)},
                {beta * t2, G(0)}},
               R, n, O(j0)));
  return PWH_OK;
}

// PulayRhoMix.cpp:5-91 (rho_old, f: mixdim columns)
int mix_pulay_core(pwh_grid* gr, int iter, const MixCols& c, int mixdim, double q0, double nel, double* d,
                   double* nel_diff) {
  const size_t n = gr->nnr;
  const double beta = 0.25;
  double* R = c.R;
  auto O = [&](int j) { return c.O[j]; };
  auto F = [&](int j) { return c.F[j]; };
  double sq = 0;
  if (iter < mixdim) {
    // PulayRhoMix.cpp:22-35
    CHK(comb_dev({{beta, R}, {1.0 - beta, O(iter - 1)}}, R, n));
    CHK(comb_dev({{1.0, R}, {-1.0, O(iter - 1)}}, F(iter), n));
    CHK(kerker_dev(gr, F(iter), 1.0, q0));
    CHK(comb_dev({{1.0, R}}, O(iter), n));
    CHK(dot_dev(F(iter), F(iter), n, &sq));
    *d = std::sqrt(sq / (double)n);
    return PWH_OK;
  }
  const int jc = iter % mixdim, jn = (iter + 1) % mixdim;
  const int m = std::min(iter + 1, mixdim) + 1;
  CHK(comb_dev({{1.0, R}, {-1.0, O(jc)}}, F(jc), n));
  CHK(kerker_dev(gr, F(jc), 1.0, q0));
  CHK(dot_dev(F(jc), F(jc), n, &sq));
  *d = std::sqrt(sq / (double)n);
  // the bordered Gram system (:54-71), symmetric from its upper triangle
  std::vector<double> A((size_t)m * m, 0.0), b(m, 0.0);
  for (int i = 0; i < m - 1; ++i) {
    for (int j = i; j < m - 1; ++j) {
      double v = 0;
      CHK(dot_dev(F(i), F(j), n, &v));
      A[i + (size_t)j * m] = v;
      A[j + (size_t)i * m] = v;
    }
    A[i + (size_t)(m - 1) * m] = 1.0;
    A[(m - 1) + (size_t)i * m] = 1.0;
  }
  b[m - 1] = 1.0;
  CHK(lu_solve_dev(A, b, m));
  // rhoe = sum_j alpha_j (rho_old_j + f_j), three histories per pass
  CHK(comb_dev({}, R, n));
  for (int j0 = 0; j0 < m - 1; j0 += 2) {
    if (j0 + 1 < m - 1)
      CHK(comb_dev({{1.0, R}, {b[j0], O(j0)}, {b[j0], F(j0)}, {b[j0 + 1], O(j0 + 1)}, {b[j0 + 1], F(j0 + 1)}},
                   R, n));
    else
      CHK(comb_dev({{1.0, R}, {b[j0], O(j0)}, {b[j0], F(j0)}}, R, n));
  }
  CHK(normalize_dev(R, n, gr->omega, nel, nel_diff));
  CHK(comb_dev({{1.0, R}}, O(jn), n));
  return PWH_OK;
}

// BroydenRhoMix.cpp:33-103 (rho_old, f: 2 columns; df, u: mixdim; a [mixdim x mixdim] host)
int mix_broyden_core(int iter, double alpha, double w0, const MixCols& c, double* a, int mixdim, size_t n,
                     double* d) {
  double* R = c.R;
  auto O = [&](int j) { return c.O[j]; };
  auto F = [&](int j) { return c.F[j]; };
  auto DF = [&](int j) { return c.DF[j]; };
  auto U = [&](int j) { return c.U[j]; };
  const int m = std::min(iter + 1, mixdim);
  const int jc = iter % m, kp = (iter - 1) % 2, kc = iter % 2;
  double sq = 0;
  CHK(comb_dev({{1.0, R}, {-1.0, O(kp)}}, F(kc), n, nullptr, R, O(kp), &sq));
  *d = std::sqrt(sq / (double)n);
  CHK(comb_dev({{1.0, F(kc)}, {-1.0, F(kp)}}, DF(jc), n));
  double t1 = 0;
  CHK(dot_dev(DF(jc), DF(jc), n, &t1));
  t1 = std::sqrt(t1);
  if (t1 > 1.e-8) t1 = 1.0 / t1;
  CHK(comb_dev({{t1, DF(jc)}}, DF(jc), n));
  CHK(comb_dev({{alpha, DF(jc)}, {t1, O(kp)}, {-t1, O(kc)}}, U(jc), n));
  std::vector<double> cv(m);
  for (int k = 0; k < m; ++k) CHK(dot_dev(DF(k), F(kc), n, &cv[k]));
  for (int k = 0; k < m; ++k) {
    double v = 0;
    CHK(dot_dev(DF(jc), DF(k), n, &v));
    a[k + (size_t)jc * mixdim] = v;
    a[jc + (size_t)k * mixdim] = v;
  }
  std::vector<double> beta(a, a + (size_t)mixdim * mixdim);
  for (int k = 0; k < m; ++k) beta[k + (size_t)k * mixdim] += w0 * w0;
  CHK(lu_inverse_dev(beta, m, mixdim));
  std::vector<double> gamma(m, 0.0);
  for (int l = 0; l < m; ++l)
    for (int k = 0; k < m; ++k) gamma[l] = gamma[l] + cv[k] * beta[k + (size_t)l * mixdim];
  CHK(comb_dev({{1.0, O(kp)}, {alpha, F(kc)}}, R, n));
  for (int l = 0; l < m; ++l) CHK(comb_dev({{1.0, R}, {-gamma[l], U(l)}}, R, n));
  CHK(comb_dev({{1.0, R}}, O(kc), n));
  return PWH_OK;
}

// host columns of a reference history array <-> device work vectors v[first ...]
int cols_up(int first, const double* h, int ncols, size_t n) {
  ScfWork& w = scfw();
  for (int k = 0; k < ncols; ++k) CHK(scf_up(w.v[first + k], h + k * n, n));
  return PWH_OK;
}
int cols_down(double* h, int first, int ncols, size_t n) {
  ScfWork& w = scfw();
  for (int k = 0; k < ncols; ++k) CHK(scf_down(h + k * n, w.v[first + k].p, n));
  return PWH_OK;
}

}  // namespace

extern "C" {

int pwh_mix_simple(double beta, double* rhoe, double* rho_old, int nnr, double* d) {
  if (!rhoe || !rho_old || !d || nnr < 1) return fail(PWH_ERR_ARG, "bad mix arguments");
  CHK(ensure_init());
  ScfWork& w = scfw();
  const size_t n = nnr;
  CHK(scf_up(w.v[0], rhoe, n));
  CHK(scf_up(w.v[1], rho_old, n));
  MixCols c;
  c.R = w.v[0].p;
  c.O[0] = w.v[1].p;
  CHK(mix_simple_core(beta, c, n, d));
  CHK(scf_down(rhoe, w.v[0].p, n));
  CHK(scf_down(rho_old, w.v[1].p, n));
  return scf_sync();
}

int pwh_mix_kerker(pwh_grid* gr, double beta, double q0, double* rhoe, double* rho_old, double* d) {
  if (!gr || !rhoe || !rho_old || !d) return fail(PWH_ERR_ARG, "bad mix arguments");
  ScfWork& w = scfw();
  const size_t n = gr->nnr;
  CHK(scf_up(w.v[0], rhoe, n));
  CHK(scf_up(w.v[1], rho_old, n));
  CHK(w.v[2].ensure(n));
  MixCols c;
  c.R = w.v[0].p;
  c.O[0] = w.v[1].p;
  c.G[0] = w.v[2].p;
  CHK(mix_kerker_core(gr, beta, q0, c, d));
  CHK(scf_down(rhoe, w.v[0].p, n));
  CHK(scf_down(rho_old, w.v[1].p, n));
  return scf_sync();
}

int pwh_kerker_prec(pwh_grid* gr, double* diff_rhoe, int idx1, double beta, double q0) {
  if (!gr || !diff_rhoe || idx1 < 1) return fail(PWH_ERR_ARG, "bad kerker_prec arguments");
  ScfWork& w = scfw();
  const size_t n = gr->nnr;
  double* col = diff_rhoe + (size_t)(idx1 - 1) * n;
  CHK(scf_up(w.v[0], col, n));
  CHK(kerker_dev(gr, w.v[0].p, beta, q0));
  CHK(scf_down(col, w.v[0].p, n));
  return scf_sync();
}

int pwh_mix_anderson(int iter, double beta, double* rhoe, double* rho_old, double* f, double* gv, int nnr,
                     double* d) {
  if (!rhoe || !rho_old || !f || !gv || !d || nnr < 1 || iter < 1) return fail(PWH_ERR_ARG, "bad mix arguments");
  CHK(ensure_init());
  ScfWork& w = scfw();
  const size_t n = nnr;
  // device: 0 rhoe, 1..3 rho_old columns, 4..6 f columns, 7..8 g columns
  CHK(scf_up(w.v[0], rhoe, n));
  CHK(cols_up(1, rho_old, 3, n));
  CHK(cols_up(4, f, 3, n));
  CHK(cols_up(7, gv, 2, n));
  MixCols c;
  c.R = w.v[0].p;
  for (int k = 0; k < 3; ++k) { c.O[k] = w.v[1 + k].p; c.F[k] = w.v[4 + k].p; }
  for (int k = 0; k < 2; ++k) c.G[k] = w.v[7 + k].p;
  CHK(mix_anderson_core(iter, beta, c, n, d));
  CHK(scf_down(rhoe, c.R, n));
  CHK(cols_down(rho_old, 1, 3, n));
  CHK(cols_down(f, 4, 3, n));
  CHK(cols_down(gv, 7, 2, n));
  return scf_sync();
}

int pwh_mix_pulay(pwh_grid* gr, int iter, double* rhoe, double* rho_old, double* f, int mixdim, double q0,
                  double nel, double* d, double* nel_diff) {
  if (!gr || !rhoe || !rho_old || !f || !d || iter < 1) return fail(PWH_ERR_ARG, "bad mix arguments");
  if (mixdim < 2 || mixdim > 7) return fail(PWH_ERR_ARG, "PulayRhoMix: mixdim must be in 2..7");
  ScfWork& w = scfw();
  const size_t n = gr->nnr;
  // device: 0 rhoe, 1..mixdim rho_old columns, 8..7+mixdim f columns
  CHK(scf_up(w.v[0], rhoe, n));
  CHK(cols_up(1, rho_old, mixdim, n));
  CHK(cols_up(8, f, mixdim, n));
  MixCols c;
  c.R = w.v[0].p;
  for (int k = 0; k < mixdim; ++k) { c.O[k] = w.v[1 + k].p; c.F[k] = w.v[8 + k].p; }
  CHK(mix_pulay_core(gr, iter, c, mixdim, q0, nel, d, nel_diff));
  CHK(scf_down(rhoe, c.R, n));
  CHK(cols_down(rho_old, 1, mixdim, n));
  CHK(cols_down(f, 8, mixdim, n));
  return scf_sync();
}

int pwh_mix_broyden(int iter, double alpha, double w0, double* rho, double* rho_old, double* f, double* df,
                    double* u, double* a, int mixdim, int nnr, double* d) {
  if (!rho || !rho_old || !f || !df || !u || !a || !d || nnr < 1 || iter < 1)
    return fail(PWH_ERR_ARG, "bad mix arguments");
  if (mixdim < 2 || mixdim > 8) return fail(PWH_ERR_ARG, "BroydenRhoMix: mixdim must be in 2..8");
  CHK(ensure_init());
  ScfWork& w = scfw();
  const size_t n = nnr;
  // device: 0 rho, 1..2 rho_old, 3..4 f, 5..4+mixdim df, 13..12+mixdim u
  CHK(scf_up(w.v[0], rho, n));
  CHK(cols_up(1, rho_old, 2, n));
  CHK(cols_up(3, f, 2, n));
  CHK(cols_up(5, df, mixdim, n));
  CHK(cols_up(13, u, mixdim, n));
  MixCols c;
  c.R = w.v[0].p;
  for (int k = 0; k < 2; ++k) { c.O[k] = w.v[1 + k].p; c.F[k] = w.v[3 + k].p; }
  for (int k = 0; k < mixdim; ++k) { c.DF[k] = w.v[5 + k].p; c.U[k] = w.v[13 + k].p; }
  CHK(mix_broyden_core(iter, alpha, w0, c, a, mixdim, n, d));
  CHK(scf_down(rho, c.R, n));
  CHK(cols_down(rho_old, 1, 2, n));
  CHK(cols_down(f, 3, 2, n));
  CHK(cols_down(df, 5, mixdim, n));
  CHK(cols_down(u, 13, mixdim, n));
  return scf_sync();
}

// ---------------------------------------------------------------------------
// The density side of an SCF loop held in HBM (density, V_eff, mixing history)
// ---------------------------------------------------------------------------
}  // extern "C"

struct pwh_scf_state {
  pwh_grid* grid = nullptr;
  int mixer = 0, mixdim = 0, ncols = 0;
  Buf<double> rho, veff, cols;   // cols: the history columns, nnr each
  std::vector<double> a;         // Broyden's a [mixdim x mixdim]
  MixCols mc;                    // R = rho, the history columns in cols
};

extern "C" {

int pwh_scf_state_create(pwh_grid* gr, int mixer, int mixdim, pwh_scf_state** out) {
  if (!gr || !out) return fail(PWH_ERR_ARG, "null argument");
  if (mixer != 0 && mixer != 1 && mixer != 2 && mixer != 4)
    return fail(PWH_ERR_UNSUPPORTED, "scf state mixers: 0 simple, 1 Kerker, 2 Anderson, 4 Broyden");
  if (mixer == 4 && (mixdim < 2 || mixdim > 8)) return fail(PWH_ERR_ARG, "BroydenRhoMix: mixdim must be in 2..8");
  CHK(ensure_init());
  std::unique_ptr<pwh_scf_state> st(new pwh_scf_state());
  st->grid = gr;
  st->mixer = mixer;
  st->mixdim = mixer == 4 ? mixdim : 0;
  st->ncols = mixer == 0 ? 1 : mixer == 1 ? 2 : mixer == 2 ? 8 : 4 + 2 * mixdim;
  const size_t n = gr->nnr;
  CHK(st->rho.ensure(n));
  CHK(st->veff.ensure(n));
  CHK(st->cols.ensure(n * st->ncols));
  MixCols& c = st->mc;
  c.R = st->rho.p;
  auto col = [&](int k) { return st->cols.p + (size_t)k * n; };
  if (mixer == 0) {
    c.O[0] = col(0);
  } else if (mixer == 1) {
    c.O[0] = col(0);
    c.G[0] = col(1);
  } else if (mixer == 2) {
    for (int k = 0; k < 3; ++k) { c.O[k] = col(k); c.F[k] = col(3 + k); }
    for (int k = 0; k < 2; ++k) c.G[k] = col(6 + k);
  } else {
    for (int k = 0; k < 2; ++k) { c.O[k] = col(k); c.F[k] = col(2 + k); }
    for (int k = 0; k < mixdim; ++k) { c.DF[k] = col(4 + k); c.U[k] = col(4 + mixdim + k); }
    st->a.assign((size_t)mixdim * mixdim, 0.0);
  }
  *out = st.release();
  return PWH_OK;
}

int pwh_scf_state_destroy(pwh_scf_state* st) {
  if (!st) return PWH_OK;
  if (g.stream) hipStreamSynchronize(g.stream);
  delete st;
  return PWH_OK;
}

int pwh_scf_state_set_rho(pwh_scf_state* st, const double* rho) {
  if (!st || !rho) return fail(PWH_ERR_ARG, "null argument");
  const size_t n = st->grid->nnr;
  CHK(upload(st->rho, rho, n));
  HIPCHK(fill_zero(st->cols.p, n * st->ncols, g.stream));
  // the reference's history arrays before the first mixing step (ChebySCFLoop.cpp:121-168):
  // rho_old = the start density for the simple and Kerker mixers; the Anderson and Broyden
  // histories start at zero
  if (st->mixer == 0 || st->mixer == 1)
    HIPCHK(hipMemcpyAsync(st->mc.O[0], st->rho.p, sizeof(double) * n, hipMemcpyDeviceToDevice, g.stream));
  std::fill(st->a.begin(), st->a.end(), 0.0);
  return scf_sync();
}

int pwh_scf_state_get(pwh_scf_state* st, double* rho, double* veff) {
  if (!st) return fail(PWH_ERR_ARG, "null argument");
  const size_t n = st->grid->nnr;
  if (rho) CHK(scf_down(rho, st->rho.p, n));
  if (veff) CHK(scf_down(veff, st->veff.p, n));
  return scf_sync();
}

int pwh_scf_state_mix(pwh_scf_state* st, int iter, double beta, double q0, double alpha, double w0, double* d) {
  if (!st || !d || iter < 1) return fail(PWH_ERR_ARG, "bad mix arguments");
  const size_t n = st->grid->nnr;
  switch (st->mixer) {
    case 0: CHK(mix_simple_core(beta, st->mc, n, d)); break;
    case 1: CHK(mix_kerker_core(st->grid, beta, q0, st->mc, d)); break;
    case 2: CHK(mix_anderson_core(iter, beta, st->mc, n, d)); break;
    default: CHK(mix_broyden_core(iter, alpha, w0, st->mc, st->a.data(), st->mixdim, n, d)); break;
  }
  return scf_sync();
}

int pwh_scf_state_energy(pwh_scf_state* st, pwh_energy* e) {
  if (!st || !e) return fail(PWH_ERR_ARG, "null argument");
  return eval_energy_core(st->grid, st->rho.p, st->veff.p, e);
}

int pwh_scf_state_density(pwh_system* s, pwh_scf_state* st, int nstates, const double* w, double omega,
                          double* ekin, double* charge) {
  CHK(check_ld(s, nstates));
  if (!st || !w || !ekin || !charge || !(omega > 0.0)) return fail(PWH_ERR_ARG, "bad scf_state_density arguments");
  if ((size_t)s->nnr != st->grid->nnr) return fail(PWH_ERR_ARG, "the grid and the system differ in NNR");
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  const size_t n = (size_t)nstates * s->ngw_pad;
  CHK(s->dtmp.ensure(n));
  HIPCHK(hipMemcpyAsync(s->dtmp.p, s->X.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, g.stream));
  return eval_rho_core(s, nullptr, s->dtmp.p, nstates, w, omega, nullptr, ekin, charge, st->rho.p);
}

int pwh_scf_state_set_vloc(pwh_system* s, pwh_scf_state* st) {
  if (!s || !st) return fail(PWH_ERR_ARG, "null argument");
  if ((size_t)s->nnr != st->grid->nnr) return fail(PWH_ERR_ARG, "the grid and the system differ in NNR");
  const size_t n = (size_t)s->nnr;
  CHK(s->V64.ensure(n));
  CHK(s->V32.ensure(n));
  HIPCHK(hipMemcpyAsync(s->V64.p, st->veff.p, sizeof(double) * n, hipMemcpyDeviceToDevice, g.stream));
  HIPCHK(to_float(st->veff.p, n, s->V32.p, g.stream));
  s->vset = true;
  return PWH_OK;
}

// ENL (EvalENL.cpp:5-49) of nb states held in sphere order in Xs (read only)
static int eval_enl_core(pwh_system* s, const cx<double>* Xs, int nb, const double* focc, double wkpt,
                         double* fnl, double* enl) {
  const int ldp = s->ngw_pad;
  CHK(upload(s->rho_w, focc, nb));
  const int nproj = s->poff.empty() ? 0 : s->poff.back() + s->na.back() * s->lm.back();
  std::vector<double> rows(std::max(nproj, 1), 0.0);
  CHK(s->Q64.ensure((size_t)std::max(nproj, 1) * nb));
  CHK(s->rho_k.ensure(std::max(nproj, 1)));
  // FNL = beta^H C0 per species (EvalENL.cpp:28-34), then sum_ist FOCC |FNL|^2 per row
  for (int sp = 0; sp < s->nsp; ++sp) {
    if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
    const int M = s->na[sp] * s->lm[sp];
    const size_t pe = proj_part_elems(M, nb, s->ngw);
    CHK(s->part64.ensure(pe));
    BetaGen bg{s->Bt64.p + (size_t)s->poff[sp] * ldp, M, s->lm[sp], ldp};
    HIPCHK(gemm_proj<double>(bg, Xs, ldp, nb, s->ngw, s->Q64.p + (size_t)s->poff[sp] * nb, s->part64.p, pe,
                             PREC_FP64, g.stream));
  }
  if (nproj > 0) {
    HIPCHK(enl_rows(s->Q64.p, nb, nproj, nb, s->rho_w.p, s->rho_k.p, g.stream));
    HIPCHK(hipMemcpyAsync(rows.data(), s->rho_k.p, sizeof(double) * nproj, hipMemcpyDeviceToHost, g.stream));
    if (fnl)
      HIPCHK(hipMemcpyAsync(fnl, s->Q64.p, sizeof(cx<double>) * nproj * nb, hipMemcpyDeviceToHost, g.stream));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  // ENL = sum over (species, atom, lm) in the reference's loop order of WKPT enl_m WNL
  double e = 0.0;
  for (int sp = 0; sp < s->nsp; ++sp) {
    if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
    for (int a = 0; a < s->na[sp]; ++a)
      for (int l = 0; l < s->lm[sp]; ++l)
        e = e + wkpt * rows[s->poff[sp] + a * s->lm[sp] + l] * s->wl[4 * sp + l];
  }
  *enl = e;
  return PWH_OK;
}

int pwh_eval_enl(pwh_system* s, const double* c0, int ld, int nstates, const double* focc, double wkpt,
                 double* fnl, double* enl) {
  CHK(check_ld(s, nstates));
  if (!c0 || !focc || !enl || ld < s->ngw) return fail(PWH_ERR_ARG, "bad eval_enl arguments");
  const int nb = nstates, ldp = s->ngw_pad;
  CHK(s->stage.ensure((size_t)s->ngw * nb));
  HIPCHK(hipMemcpy2DAsync(s->stage.p, sizeof(cx<double>) * s->ngw, c0, sizeof(cx<double>) * ld,
                          sizeof(cx<double>) * s->ngw, nb, hipMemcpyHostToDevice, g.stream));
  CHK(s->X.ensure((size_t)nb * ldp));
  HIPCHK(permute_in_f64(s->stage.p, s->ngw, s->d_perm.p, s->ngw, ldp, s->X.p, nb, g.stream));
  return eval_enl_core(s, s->X.p, nb, focc, wkpt, fnl, enl);
}

int pwh_resident_eval_enl(pwh_system* s, int nstates, const double* focc, double wkpt, double* fnl, double* enl) {
  CHK(check_ld(s, nstates));
  if (!focc || !enl) return fail(PWH_ERR_ARG, "bad eval_enl arguments");
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  return eval_enl_core(s, s->X.p, nstates, focc, wkpt, fnl, enl);
}

// ---------------------------------------------------------------------------
// Set-up on the GPU: PhaseFactor, StructureFactor, NLFormFactor, FormFactor
// ---------------------------------------------------------------------------
}  // extern "C"

namespace pwh {
void reduced_coords(const double A[9], const double* tau, int n, double* out);
}

namespace {

struct EiTables {
  Buf<cx<double>> ei;
  Buf<double> taup;
  Buf<int> na;
  int ldt = 0;
  int idx0[3] = {0, 0, 0};
};

// EI tables of every (species, atom) slot; atoms past NA(is) keep zeros
int ei_build(EiTables& t, const double A[9], const int nr[3], int nsp, const int* na, int nax, const double* tau) {
  const int slots = nsp * nax;
  std::vector<double> tp((size_t)3 * slots, 0.0);
  for (int is = 0; is < nsp; ++is) {
    if (na[is] < 0 || na[is] > nax) return fail(PWH_ERR_ARG, "NA(is) outside 0..NAX");
    pwh::reduced_coords(A, tau + (size_t)3 * nax * is, na[is], tp.data() + (size_t)3 * nax * is);
  }
  t.ldt = std::max(nr[0], std::max(nr[1], nr[2])) + 2;
  for (int d = 0; d < 3; ++d) t.idx0[d] = (nr[d] + 1) / 2;
  CHK(t.ei.ensure((size_t)slots * 3 * t.ldt));
  HIPCHK(hipMemsetAsync(t.ei.p, 0, sizeof(cx<double>) * slots * 3 * t.ldt, g.stream));
  CHK(upload(t.taup, tp.data(), tp.size()));
  CHK(upload(t.na, na, nsp));
  // the tables of every slot are written; empty slots are overwritten by zeros below
  HIPCHK(ei_tables(t.taup.p, slots, nr, t.ldt, t.ei.p, g.stream));
  for (int is = 0; is < nsp; ++is)
    if (na[is] < nax)
      HIPCHK(hipMemsetAsync(t.ei.p + (size_t)(is * nax + na[is]) * 3 * t.ldt, 0,
                            sizeof(cx<double>) * (nax - na[is]) * 3 * t.ldt, g.stream));
  return PWH_OK;
}

int pseudo_check(const pwh_pseudo_desc* ps) {
  if (!ps || ps->nsp < 1 || !ps->mmax || !ps->clog || !ps->r || !ps->vion || !ps->l_loc || ps->maxr < 1 ||
      ps->nang < 1)
    return fail(PWH_ERR_ARG, "bad pseudopotential description");
  for (int is = 0; is < ps->nsp; ++is)
    if (ps->mmax[is] < 3 || ps->mmax[is] > ps->maxr || ps->mmax[is] % 2 == 0)
      return fail(PWH_ERR_ARG, "MMAX must be odd and <= MAXR (simpson.cpp:8-11)");
  return PWH_OK;
}

}  // namespace

extern "C" {

int pwh_projector_mode(pwh_system* s, int mode, const double A[9], const double* tau) {
  if (!s || (mode != 0 && mode != 1)) return fail(PWH_ERR_ARG, "bad projector mode");
  CHK(ensure_init());
  if (mode == 0) {
    s->beta_fly = 0;
    return PWH_OK;
  }
  if (!A || !tau) return fail(PWH_ERR_ARG, "the on-the-fly projectors need the cell and the atoms");
  if (s->ldt > 1024 + 2) return fail(PWH_ERR_UNSUPPORTED, "grid too large for the packed table indices");
  EiTables t;
  CHK(ei_build(t, A, s->nr, s->nsp, s->na.data(), s->nax, tau));
  const size_t n = (size_t)s->nsp * s->nax * 3 * t.ldt;
  if (t.ldt != s->ldt) return fail(PWH_ERR_ARG, "phase table length mismatch");
  CHK(s->ei32.ensure(n));
  HIPCHK((convert_cx<double, float>(t.ei.p, s->ei32.p, n, g.stream)));
  CHK(upload(s->pkg32, s->pkg_host.data(), s->pkg_host.size()));
  HIPCHK(hipStreamSynchronize(g.stream));
  s->beta_fly = 1;
  ++g.layout_gen;
  return PWH_OK;
}

int pwh_phase_factor_gpu(const double A[9], const int nr[3], int nsp, const int* na, int nax, const double* tau,
                         int nkpt, const int* ngw, int ngwx, const int* igk, int ngx, const int* in123,
                         double* eigr) {
  if (!A || !nr || nsp < 1 || !na || nax < 1 || !tau || nkpt < 1 || !ngw || !igk || !in123 || !eigr || ngwx < 1)
    return fail(PWH_ERR_ARG, "bad phase_factor arguments");
  CHK(ensure_init());
  EiTables t;
  CHK(ei_build(t, A, nr, nsp, na, nax, tau));
  Buf<int> digk, din;
  Buf<cx<double>> de;
  CHK(upload(din, in123, (size_t)3 * ngx));
  CHK(digk.ensure((size_t)ngwx));
  const size_t per_k = (size_t)ngwx * nax * nsp;
  CHK(de.ensure(per_k));
  for (int ik = 0; ik < nkpt; ++ik) {
    if (ngw[ik] < 0 || ngw[ik] > ngwx) return fail(PWH_ERR_ARG, "NGW(ik) outside 0..NGWX");
    for (int ig = 0; ig < ngw[ik]; ++ig)
      if (igk[ig + (size_t)ngwx * ik] < 1 || igk[ig + (size_t)ngwx * ik] > ngx)
        return fail(PWH_ERR_ARG, "IGK outside 1..NGX");
    HIPCHK(hipMemcpyAsync(digk.p, igk + (size_t)ngwx * ik, sizeof(int) * ngwx, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemsetAsync(de.p, 0, sizeof(cx<double>) * per_k, g.stream));
    if (ngw[ik] > 0)
      HIPCHK(eigr_dev(t.ei.p, t.ldt, t.na.p, nsp, nax, digk.p, din.p, ngx, ngw[ik], ngwx, t.idx0, de.p, g.stream));
    HIPCHK(hipMemcpyAsync(eigr + 2 * per_k * ik, de.p, sizeof(cx<double>) * per_k, hipMemcpyDeviceToHost,
                          g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
  }
  return PWH_OK;
}

int pwh_structure_factor(const double A[9], const int nr[3], int nsp, const int* na, int nax, const double* tau,
                         int ng, int ngx, const int* in123, double* sfac) {
  if (!A || !nr || nsp < 1 || !na || nax < 1 || !tau || ng < 1 || ngx < ng || !in123 || !sfac)
    return fail(PWH_ERR_ARG, "bad structure_factor arguments");
  CHK(ensure_init());
  EiTables t;
  CHK(ei_build(t, A, nr, nsp, na, nax, tau));
  Buf<int> din;
  Buf<cx<double>> ds;
  CHK(upload(din, in123, (size_t)3 * ngx));
  CHK(ds.ensure((size_t)nsp * ng));
  HIPCHK(sfac_dev(t.ei.p, t.ldt, t.na.p, nax, nsp, din.p, ngx, ng, t.idx0, ds.p, g.stream));
  HIPCHK(hipMemcpyAsync(sfac, ds.p, sizeof(cx<double>) * nsp * ng, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_nl_form_factor(const pwh_pseudo_desc* ps, double omega, double tpiba, int nkpt, const int* ngw, int ngwx,
                       const double* xkg, const double* ggk, const double* xk, int nlmax, double* wnl, double* pkg,
                       double* pkg_a) {
  CHK(pseudo_check(ps));
  if (!ps->psi || !ps->l_max || nkpt < 1 || !ngw || ngwx < 1 || !xkg || !ggk || !xk || nlmax < 1 || !wnl ||
      !pkg_a || !(omega > 0))
    return fail(PWH_ERR_ARG, "bad nl_form_factor arguments");
  CHK(ensure_init());
  const int nsp = ps->nsp, maxr = ps->maxr;
  for (int is = 0; is < nsp; ++is) {
    int ml = 0;
    for (int l = 1; l <= ps->l_max[is]; ++l) {
      if (l == ps->l_loc[is]) continue;
      if (l >= 4) return fail(PWH_ERR_UNSUPPORTED, "not implemented i_l = " + std::to_string(l));
      if (l > ps->nang || ps->l_loc[is] < 1 || ps->l_loc[is] > ps->nang)
        return fail(PWH_ERR_ARG, "channel outside the pseudopotential data");
      ml += 2 * l - 1;
    }
    if (ml > nlmax) return fail(PWH_ERR_ARG, "more projector components than NLMAX");
  }
  const double fpibo = 4.0 * M_PI / omega;
  Buf<double> dr, dfun, dfun1, dvl, dvloc, dpsi, dxkg, dggk, dw;
  Buf<double> dpkg;
  CHK(dr.ensure(maxr));
  CHK(dfun.ensure(maxr));
  CHK(dfun1.ensure(maxr));
  CHK(dw.ensure(1));
  CHK(dxkg.ensure(ngwx));
  CHK(dggk.ensure((size_t)3 * ngwx));
  const size_t per_k = (size_t)nlmax * ngwx * nsp;
  CHK(dpkg.ensure(per_k));
  std::vector<double> hp(per_k);
  for (int ik = 0; ik < nkpt; ++ik) {
    if (ngw[ik] < 0 || ngw[ik] > ngwx) return fail(PWH_ERR_ARG, "NGW(ik) outside 0..NGWX");
    HIPCHK(hipMemcpyAsync(dxkg.p, xkg + (size_t)ngwx * ik, sizeof(double) * ngwx, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(dggk.p, ggk + (size_t)3 * ngwx * ik, sizeof(double) * 3 * ngwx, hipMemcpyHostToDevice,
                          g.stream));
    // start from the caller's PKG_A so components no channel writes keep their values
    HIPCHK(hipMemcpyAsync(dpkg.p, pkg_a + per_k * ik, sizeof(double) * per_k, hipMemcpyHostToDevice, g.stream));
    for (int is = 0; is < nsp; ++is) {
      const int m = ps->mmax[is];
      const double h = ps->clog[is];
      CHK(upload(dr, ps->r + (size_t)maxr * is, m));
      CHK(upload(dvloc, ps->vion + (size_t)maxr * (is + (size_t)nsp * (ps->l_loc[is] - 1)), m));
      int ml = 0;
      for (int l = 1; l <= ps->l_max[is]; ++l) {
        if (l == ps->l_loc[is]) continue;
        CHK(upload(dvl, ps->vion + (size_t)maxr * (is + (size_t)nsp * (l - 1)), m));
        CHK(upload(dpsi, ps->psi + (size_t)maxr * (is + (size_t)nsp * (l - 1)), m));
        HIPCHK(nl_prep(dr.p, dvl.p, dvloc.p, dpsi.p, m, dfun.p, dfun1.p, g.stream));
        if (ik == 0) {
          HIPCHK(nl_wnl(dfun1.p, m, h, fpibo, l, dw.p, g.stream));
          double w = 0;
          HIPCHK(hipMemcpyAsync(&w, dw.p, sizeof(double), hipMemcpyDeviceToHost, g.stream));
          HIPCHK(hipStreamSynchronize(g.stream));
          for (int j = 0; j < 2 * l - 1; ++j) wnl[is + (size_t)nsp * (ml + j)] = w;
        }
        if (ngw[ik] > 0)
          HIPCHK(nl_pkg(dr.p, dfun.p, m, h, l, ml, dxkg.p, dggk.p, xk + 3 * ik, ngw[ik], tpiba, nlmax,
                        dpkg.p + (size_t)nlmax * ngwx * is, g.stream));
        ml += 2 * l - 1;
      }
    }
    HIPCHK(hipMemcpyAsync(hp.data(), dpkg.p, sizeof(double) * per_k, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    for (size_t i = 0; i < per_k; ++i) pkg_a[per_k * ik + i] = hp[i];
    if (pkg) {
      // PKG(ig, is, ik, m) = PKG_A(m, ig, is, ik) for the written components (:119-121)
      for (int is = 0; is < nsp; ++is) {
        int ml = 0;
        for (int l = 1; l <= ps->l_max[is]; ++l) {
          if (l == ps->l_loc[is]) continue;
          for (int j = 0; j < 2 * l - 1; ++j)
            for (int ig = 0; ig < ngw[ik]; ++ig)
              pkg[ig + (size_t)ngwx * (is + (size_t)nsp * (ik + (size_t)nkpt * (ml + j)))] =
                  hp[(ml + j) + (size_t)nlmax * (ig + (size_t)ngwx * is)];
          ml += 2 * l - 1;
        }
      }
    }
  }
  return PWH_OK;
}

int pwh_form_factor(const pwh_pseudo_desc* ps, const int* na, double omega, double tpiba2, int ng, const double* g2,
                    double* vps, double* rhops, double* eself) {
  CHK(pseudo_check(ps));
  if (!na || ng < 1 || !g2 || !vps || !rhops || !eself || !ps->zv || !ps->rgauss || !(omega > 0) || !(tpiba2 > 0))
    return fail(PWH_ERR_ARG, "bad form_factor arguments");
  for (int is = 0; is < ps->nsp; ++is)
    if (ps->l_loc[is] < 1 || ps->l_loc[is] > ps->nang) return fail(PWH_ERR_ARG, "ERROR in formf: ll < 1");
  CHK(ensure_init());
  const int nsp = ps->nsp, maxr = ps->maxr;
  // ESELF (FormFactor.cpp:19-24)
  double es = 0.0;
  for (int is = 0; is < nsp; ++is) es = es + ps->zv[is] * ps->zv[is] / ps->rgauss[is] * na[is];
  *eself = es * (1.0 / std::sqrt(2.0 * M_PI));
  const double tpiba = std::sqrt(tpiba2), fpibom = 4.0 * M_PI / omega;
  Buf<double> dr, dvl, dvloc, dg, dv, drh;
  CHK(upload(dg, g2, ng));
  CHK(upload(dv, vps, (size_t)nsp * ng));
  CHK(upload(drh, rhops, (size_t)nsp * ng));
  for (int is = 0; is < nsp; ++is) {
    const int m = ps->mmax[is];
    CHK(upload(dr, ps->r + (size_t)maxr * is, m));
    CHK(upload(dvl, ps->vion + (size_t)maxr * (is + (size_t)nsp * (ps->l_loc[is] - 1)), m));
    CHK(dvloc.ensure(m));
    HIPCHK(ff_prep(dr.p, dvl.p, m, ps->zv[is], ps->rgauss[is], dvloc.p, g.stream));
    const double facc = 0.25 * tpiba2 * ps->rgauss[is] * ps->rgauss[is];
    HIPCHK(ff_vps(dr.p, dvloc.p, m, ps->clog[is], dg.p, ng, tpiba, fpibom, ps->zv[is], omega, facc, nsp, is, dv.p,
                  drh.p, g.stream));
  }
  HIPCHK(hipMemcpyAsync(vps, dv.p, sizeof(double) * nsp * ng, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipMemcpyAsync(rhops, drh.p, sizeof(double) * nsp * ng, hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_resident_set(pwh_system* s, const double* X, int nstates) {
  CHK(check_ld(s, nstates));
  CHK(load_block(s, X, nstates));
  const size_t n = (size_t)nstates * s->ngw_pad;
  CHK(s->X32.ensure(n));
  HIPCHK((convert_cx<double, float>(s->X.p, s->X32.p, n, g.stream)));
  s->resident_nb = nstates;
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_resident_get(pwh_system* s, double* X, int nstates) {
  CHK(check_ld(s, nstates));
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  return store_block(s, s->X.p, X, nstates);
}

int pwh_resident_apply(pwh_system* s, int nstates, int precision) {
  CHK(check_ld(s, nstates));
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  if (precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  const size_t n = (size_t)nstates * s->ngw_pad;
  if (precision == PWH_FP64) {
    CHK(s->HX.ensure(n));
    return hpsi<double>(s, s->X.p, s->HX.p, nstates, 1.0, 0.0, 0.0, nullptr, PREC_FP64);
  }
  // FP32 modes use the FP32 copy made by pwh_resident_set
  CHK(s->Y32.ensure(n));
  return hpsi<float>(s, s->X32.p, s->Y32.p, nstates, 1.0, 0.0, 0.0, nullptr, precision);
}

int pwh_autotune(pwh_system* s, int nstates, int precision, int* target_blocks, double* best_ms) {
  if (!s) return fail(PWH_ERR_ARG, "null system");
  if (precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  if (target_blocks && *target_blocks != 0) {
    // a target known from an earlier run: set it, time nothing
    if (*target_blocks < 64 || *target_blocks > 65536) return fail(PWH_ERR_ARG, "target outside [64, 65536]");
    if (s->tune[precision] != *target_blocks) ++g.layout_gen;
    s->tune[precision] = *target_blocks;
    if (best_ms) *best_ms = 0.0;
    return PWH_OK;
  }
  CHK(check_ld(s, nstates));
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  if (!s->vset) return fail(PWH_ERR_ARG, "local potential not set (pwh_set_vloc)");
  const bool timing = g.timing;
  g.timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PWH_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = fail(PWH_ERR_HIP, "hipEventCreate");
  // ms per H|psi> of the current settings: 2 warm calls, then 5 timed, the smaller of
  // two such passes (a slow box moment does not decide)
  auto time_apply = [&](double* out) -> int {
    double best = 1e300;
    for (int pass = 0; pass < 2; ++pass) {
      const int reps = 5;
      for (int r = 0; r < 2; ++r) CHK(pwh_resident_apply(s, nstates, precision));
      if (hipEventRecord(e0, g.stream) != hipSuccess) return fail(PWH_ERR_HIP, "hipEventRecord");
      for (int r = 0; r < reps; ++r) CHK(pwh_resident_apply(s, nstates, precision));
      float el = 0.f;
      if (hipEventRecord(e1, g.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&el, e0, e1) != hipSuccess)
        return fail(PWH_ERR_HIP, "event timing");
      best = std::min(best, (double)el / reps);
    }
    *out = best;
    return PWH_OK;
  };
  // one knob over its candidates; the default (index 0 of cand) stays unless another is
  // faster by more than 1 %; every change bumps the layout generation
  auto search = [&](int* knob, const std::vector<int>& cand, double* ms_best) -> int {
    std::vector<double> ms(cand.size(), 1e300);
    for (size_t c = 0; c < cand.size(); ++c) {
      *knob = cand[c];
      ++g.layout_gen;
      CHK(time_apply(&ms[c]));
    }
    size_t best = 0;
    for (size_t c = 1; c < cand.size(); ++c)
      if (ms[c] < 0.99 * ms[best]) best = c;
    *knob = cand[best];
    ++g.layout_gen;
    *ms_best = ms[best];
    return PWH_OK;
  };
  double ms = 0.0;
  // 1. split-K block target of the projections (one to eight blocks per CU): its best
  //    value moves with the band count, the projector count and the precision
  if (rc == PWH_OK) rc = search(&s->tune[precision], {512, 256, 1024, 2048}, &ms);
  // 2. jointly: the LDS stage depth of the projection and of the expansion GEMM (FP64 /
  //    FP32; the BF16 splits have one) and the plane-pass workgroups per CU beside the
  //    overlapping nonlocal GEMMs (no cap, 3, 2, 1; fewer plane workgroups leave LDS and wave
  //    slots to GEMM workgroups).  The two interact (a deeper stage takes more of the LDS the
  //    cap frees), so every combination is timed; the default combination stays unless
  //    another is faster by more than 1 %
  if (rc == PWH_OK) {
    std::vector<int> ks = {16};
    if (precision == PWH_FP32) ks = {16, 32, 64};
    else if (precision == PWH_FP64) ks = {16, 32};
    const std::vector<int> caps = {0, 3, 2, 1};
    int* kp = &s->gks[precision][0];
    int* ke = &s->gks[precision][1];
    int* cp = &s->pcap[precision];
    const bool depths = precision <= PWH_FP32;
    double best = 1e300;
    int bk = 0, be = 0, bc = 0;
    for (size_t ia = 0; ia < ks.size() && rc == PWH_OK; ++ia)
      for (size_t ib = 0; ib < ks.size() && rc == PWH_OK; ++ib)
        for (size_t ic = 0; ic < caps.size() && rc == PWH_OK; ++ic) {
          const int a = ks[ia], b = ks[ib], c = caps[ic];
          if (depths) {
            *kp = a;
            *ke = b;
          }
          *cp = c;
          ++g.layout_gen;
          double t = 1e300;
          if ((rc = time_apply(&t)) != PWH_OK) break;
          const bool first = best == 1e300;
          if (first || t < 0.99 * best) {
            best = t;
            bk = a;
            be = b;
            bc = c;
          }
        }
    if (depths) {
      *kp = bk;
      *ke = be;
    }
    *cp = bc;
    ++g.layout_gen;
    ms = best;
    // 3. the split-K target once more, at the chosen depths and cap (the current value first)
    if (rc == PWH_OK) {
      std::vector<int> cand = {s->tune[precision]};
      for (int t : {256, 512, 1024, 2048})
        if (t != cand[0]) cand.push_back(t);
      rc = search(&s->tune[precision], cand, &ms);
    }
  }
  // 4. the radix order of the runtime FFT plans, for grids off the compiled kernels
  if (rc == PWH_OK && s->maps.path != FFT_COMPILED) {
    int v = s->fftv;
    std::vector<double> vms(3, 1e300);
    for (int c = 0; c < 3 && rc == PWH_OK; ++c) {
      FftPlan p[3];
      bool ok = true;
      for (int d = 0; d < 3; ++d) ok = ok && fft_plan_variant(s->nr[d], c, &p[d]);
      if (!ok) continue;
      for (int d = 0; d < 3; ++d) s->maps.plan[d] = p[d];
      ++g.layout_gen;
      rc = time_apply(&vms[c]);
    }
    int best = 0;
    for (int c = 1; c < 3; ++c)
      if (vms[c] < 0.99 * vms[best]) best = c;
    v = best;
    for (int d = 0; d < 3; ++d) fft_plan_variant(s->nr[d], v, &s->maps.plan[d]);
    s->fftv = v;
    ++g.layout_gen;
    if (rc == PWH_OK) ms = vms[best];
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  g.timing = timing;
  if (rc != PWH_OK) {
    s->tune[precision] = 0;
    s->gks[precision][0] = s->gks[precision][1] = 0;
    s->pcap[precision] = 0;
    return rc;
  }
  if (target_blocks) *target_blocks = s->tune[precision];
  if (best_ms) *best_ms = ms;
  return PWH_OK;
}

int pwh_tuning_get(pwh_system* s, int precision, int* out) {
  if (!s || !out || precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad tuning query");
  out[0] = s->tune[precision] ? s->tune[precision] : 512;
  const int dflt = precision >= PWH_BF16X3 ? 32 : 16;   // the BF16 split kernels stage 32 deep
  out[1] = s->gks[precision][0] ? s->gks[precision][0] : dflt;
  out[2] = s->gks[precision][1] ? s->gks[precision][1] : dflt;
  out[3] = s->maps.path == FFT_COMPILED ? -1 : s->fftv;
  out[4] = s->pcap[precision];
  return PWH_OK;
}

int pwh_tuning_set(pwh_system* s, int precision, const int* in) {
  if (!s || !in || precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad tuning setting");
  if (in[0] != 0 && (in[0] < 64 || in[0] > 65536)) return fail(PWH_ERR_ARG, "target outside [64, 65536]");
  // stage depths: the FP64 / FP32 kernels are compiled for 16 and 32 (and 64 in FP32); the
  // BF16 split kernels have their one depth (32)
  auto depth_ok = [&](int d) {
    if (d == 0) return true;
    if (precision >= PWH_BF16X3) return d == 32;
    return d == 16 || d == 32 || (d == 64 && precision == PWH_FP32);
  };
  if (!depth_ok(in[1]) || !depth_ok(in[2])) return fail(PWH_ERR_ARG, "stage depth not compiled for this precision");
  if (in[4] < 0 || in[4] > 3) return fail(PWH_ERR_ARG, "plane workgroups per CU outside [0, 3]");
  if (s->maps.path != FFT_COMPILED && in[3] != s->fftv) {
    if (in[3] < 0 || in[3] > 2) return fail(PWH_ERR_ARG, "FFT plan variant outside [0, 2]");
    FftPlan p[3];
    for (int d = 0; d < 3; ++d)
      if (!fft_plan_variant(s->nr[d], in[3], &p[d])) return fail(PWH_ERR_ARG, "FFT plan variant not available");
    for (int d = 0; d < 3; ++d) s->maps.plan[d] = p[d];
    s->fftv = in[3];
  }
  s->tune[precision] = in[0];
  if (precision <= PWH_FP32) {
    s->gks[precision][0] = in[1];
    s->gks[precision][1] = in[2];
  }
  s->pcap[precision] = in[4];
  ++g.layout_gen;
  return PWH_OK;
}

int pwh_resident_chebyfilt(pwh_system* s, int nstates, int degree, double lb, double ub,
                           int precision) {
  CHK(check_ld(s, nstates));
  if (nstates > s->resident_nb) return fail(PWH_ERR_ARG, "more states than resident");
  if (!(ub > lb)) return fail(PWH_ERR_ARG, "chebyfilt needs ub > lb");
  return filter_prec(s, s->X.p, nstates, degree, lb, ub, precision);
}

}  // extern "C"

struct pwh_batch {
  std::vector<pwh_system*> sys;
  std::vector<int> nb;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> done;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_degree = -1, g_prec = -1, g_mode = -1;
  unsigned long long g_gen = 0;   // layout generation the graph was captured at
  std::vector<double> g_lb, g_ub;
  // grouped filter (mode 2): the launch descriptors in device memory and the partial sums
  // of every projection problem of one H|psi>
  Buf<unsigned char> gdesc;
  Buf<cx<double>> gpart64;
  Buf<cx<float>> gpart32;
};

namespace {
void batch_drop_graph(pwh_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->g_degree = -1;
}

// ---------------------------------------------------------------------------
// Grouped Chebyshev filter of a batch: the filter of chebyfilt_dev / filter_on_stream for
// every system at once, each pass of each H|psi> one launch over all systems.  Every
// system keeps its own scalars (lb, ub), buffers and split-K shapes, and the passes run the
// runtime-plan FFT bodies and the one-system GEMM bodies, so a system's result equals its
// one-at-a-time filter on the runtime FFT path bit for bit.
// ---------------------------------------------------------------------------
struct GroupLaunch {
  int kind;          // 0 convert 64->32, 1 convert 32->64, 2 copy 64, 3 proj, 4 expand, 5 fft
  int fp64;          // precision of a proj / expand / fft launch
  int pass;          // fft pass (0 zinv, 1 plane, 2 zfwd)
  size_t off;        // byte offset of the descriptor array in the device buffer
  int n, blocks, rblocks;
  size_t smem;
  int key = 0;       // fft: the kernel the systems of the launch share (group_fft_key)
};

struct GroupBuilder {
  std::vector<unsigned char> host;
  std::vector<GroupLaunch> launches;
  template <typename D>
  size_t put(const std::vector<D>& v) {
    size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + sizeof(D) * v.size());
    memcpy(host.data() + off, v.data(), sizeof(D) * v.size());
    return off;
  }
};

// One step out = alpha H X + beta X + gamma Z for every system (per-system pointers and
// scalars in the vectors), in precision T.
template <typename T>
int group_hpsi(pwh_batch* b, GroupBuilder& gb, const std::vector<const cx<T>*>& X, const std::vector<cx<T>*>& out,
               const std::vector<const cx<T>*>& Z, const std::vector<double>& alpha, const std::vector<double>& beta,
               const std::vector<double>& gamma, cx<T>* part, size_t part_cap) {
  const int n = (int)b->sys.size();
  const bool fp64 = std::is_same<T, double>::value;
  const int prec = fp64 ? PREC_FP64 : PREC_FP32;
  std::vector<GroupProj<T>> pj;
  std::vector<std::vector<GroupExpand<T>>> ex;
  std::vector<GroupHpsi<T>> fh(n);
  std::vector<bool> have_nl(n, false);
  int pblk = 0, rblk = 0;
  size_t poff = 0;
  for (int i = 0; i < n; ++i) {
    pwh_system* s = b->sys[i];
    if (!s->vset) return fail(PWH_ERR_ARG, "local potential not set (pwh_set_vloc)");
    for (int kind = 0; kind < 2; ++kind)
      if (s->gks[prec][kind] != 0 && s->gks[prec][kind] != 16)
        return fail(PWH_ERR_UNSUPPORTED, "grouped filter: a GEMM stage depth tuned away from 16");
    const int nb = b->nb[i], ld = s->ngw_pad;
    for (int sp = 0; sp < s->nsp; ++sp) {
      if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
      GroupProj<T> p{};
      p.M = s->na[sp] * s->lm[sp];
      p.N = nb;
      p.K = s->ngw;
      int blocks = 0;
      if (!proj_split_shape<T>(p.M, p.N, p.K, s->tune[prec], &p.S, &p.kchunk, &blocks))
        return fail(PWH_ERR_UNSUPPORTED, "grouped filter: the GEMM stage depth is not the default");
      p.A = Sel<T>::Bt(s) + (size_t)s->poff[sp] * ld;
      p.lda = ld;
      p.B = X[i];
      p.ldb = ld;
      p.part = part + poff;
      poff += (size_t)p.S * p.M * p.N;
      if (poff > part_cap) return fail(PWH_ERR_ARG, "grouped filter: partial-sum buffer too small");
      p.C = Sel<T>::Q(s).p + (size_t)s->poff[sp] * nb;
      p.ldc = nb;
      p.wl = Sel<T>::wl(s) + 4 * sp;
      p.lm = s->lm[sp];
      p.blk0 = pblk;
      p.rblk0 = rblk;
      pblk += blocks;
      rblk += (p.M * p.N + 255) / 256;
      pj.push_back(p);
      if ((int)ex.size() <= sp) ex.resize(sp + 1);
      GroupExpand<T> e{};
      e.A = p.A;
      e.lda = ld;
      e.Q = p.C;
      e.ldq = nb;
      e.out = Sel<T>::Hnl(s).p;
      e.ldo = ld;
      e.J = p.M;
      e.N = nb;
      e.K = s->ngw;
      e.accumulate = have_nl[i] ? 1 : 0;
      e.blk0 = 0;   // set per launch below
      ex[sp].push_back(e);
      have_nl[i] = true;
    }
    GroupHpsi<T>& f = fh[i];
    f.m = s->maps;
    f.tw1 = static_cast<const cx<T>*>(Sel<T>::tw(s, 0));
    f.tw2 = static_cast<const cx<T>*>(Sel<T>::tw(s, 1));
    f.tw3 = static_cast<const cx<T>*>(Sel<T>::tw(s, 2));
    f.V = Sel<T>::V(s);
    f.kin = Sel<T>::kin(s);
    f.X = X[i];
    f.W = Sel<T>::W(s).p;
    f.Hnl = have_nl[i] ? Sel<T>::Hnl(s).p : nullptr;
    f.Z = Z[i];
    f.out = out[i];
    f.ld = ld;
    f.nb = nb;
    f.inv_nnr = (T)(1.0 / s->nnr);
    f.alpha = (T)alpha[i];
    f.beta = (T)beta[i];
    f.gamma = (T)gamma[i];
  }
  if (!pj.empty()) {
    GroupLaunch l{3, fp64, 0, gb.put(pj), (int)pj.size(), pblk, rblk, 0};
    gb.launches.push_back(l);
    for (auto& v : ex) {
      if (v.empty()) continue;
      int eb = 0;
      for (auto& e : v) {
        e.blk0 = eb;
        eb += expand_grid_blocks(e.K, e.N);
      }
      GroupLaunch le{4, fp64, 0, gb.put(v), (int)v.size(), eb, 0, 0};
      gb.launches.push_back(le);
    }
  }
  // each FFT pass: one launch per kernel the systems run (systems of one grid share one)
  for (int pass = 0; pass < 3; ++pass) {
    std::map<int, std::vector<int>> bykey;
    std::vector<int> blocks(n), cts(n);
    std::vector<size_t> sm(n);
    for (int i = 0; i < n; ++i) {
      const int key = group_fft_key<T>(pass, b->sys[i]->maps, b->nb[i], &blocks[i], &sm[i], &cts[i]);
      if (key < 0)
        return fail(PWH_ERR_UNSUPPORTED, "grouped filter: a system's plane pass needs the split path");
      bykey[key].push_back(i);
    }
    for (const auto& kv : bykey) {
      std::vector<GroupHpsi<T>> v;
      int blk = 0;
      size_t smx = 0;
      for (int i : kv.second) {
        GroupHpsi<T> f = fh[i];
        f.ct = cts[i];
        f.blk0 = blk;
        blk += blocks[i];
        smx = std::max(smx, sm[i]);
        v.push_back(f);
      }
      GroupLaunch l{5, fp64, pass, gb.put(v), (int)v.size(), blk, 0, smx, kv.first};
      gb.launches.push_back(l);
    }
  }
  return PWH_OK;
}

// kind 0/1/2 conversion of one array per system
int group_convert(pwh_batch* b, GroupBuilder& gb, int kind, const std::vector<const void*>& in,
                  const std::vector<void*>& out) {
  std::vector<GroupConv> v(b->sys.size());
  int blk = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    const size_t n = (size_t)b->nb[i] * b->sys[i]->ngw_pad;
    v[i] = GroupConv{in[i], out[i], n, blk};
    blk += (int)((n + 255) / 256);
  }
  GroupLaunch l{kind, 0, 0, gb.put(v), (int)v.size(), blk, 0, 0};
  gb.launches.push_back(l);
  return PWH_OK;
}

// Partial-sum elements of the projections of one H|psi> over the batch in precision T.
template <typename T>
size_t group_part_elems(pwh_batch* b) {
  const int prec = std::is_same<T, double>::value ? PREC_FP64 : PREC_FP32;
  size_t tot = 0;
  for (size_t i = 0; i < b->sys.size(); ++i) {
    pwh_system* s = b->sys[i];
    for (int sp = 0; sp < s->nsp; ++sp) {
      if (s->na[sp] == 0 || s->lm[sp] == 0) continue;
      const int M = s->na[sp] * s->lm[sp];
      tot += proj_part_elems(M, b->nb[i], s->ngw, s->tune[prec]);
    }
  }
  return std::max<size_t>(tot, 1);
}

// Builds the launches of the grouped filter (chebyfilt_dev's recurrence per system; FP32
// filters end with the FP64 step of cheby_last_step_fp64 unless that is switched off).
int group_filter_build(pwh_batch* b, GroupBuilder& gb, int degree, const double* lb, const double* ub,
                       int precision) {
  const int n = (int)b->sys.size();
  for (int i = 0; i < n; ++i)
    if (b->nb[i] < 2) return fail(PWH_ERR_UNSUPPORTED, "grouped filter: a system with one band");
  std::vector<double> e(n), c(n), sig(n), sig1(n);
  for (int i = 0; i < n; ++i) {
    e[i] = (ub[i] - lb[i]) / 2.0;
    c[i] = (ub[i] + lb[i]) / 2.0;
    sig[i] = sig1[i] = e[i] / (lb[i] - ub[i]);
  }
  std::vector<double> al(n), be(n), ga(n);
  if (precision == PWH_FP64) {
    CHK(b->gpart64.ensure(group_part_elems<double>(b)));
    std::vector<cx<double>*> B0(n), B1(n), B2(n);
    for (int i = 0; i < n; ++i) {
      B0[i] = b->sys[i]->X.p;
      B1[i] = b->sys[i]->A64.p;
      B2[i] = b->sys[i]->B64.p;
    }
    std::vector<const cx<double>*> in(B0.begin(), B0.end()), Z(n, nullptr);
    for (int i = 0; i < n; ++i) { al[i] = sig1[i] / e[i]; be[i] = -c[i] * sig1[i] / e[i]; ga[i] = 0.0; }
    CHK(group_hpsi<double>(b, gb, in, B1, Z, al, be, ga, b->gpart64.p, b->gpart64.cap));
    for (int id = 2; id <= degree; ++id) {
      std::vector<const cx<double>*> yin(B1.begin(), B1.end()), zz(B0.begin(), B0.end());
      for (int i = 0; i < n; ++i) {
        const double s2 = 1.0 / (2.0 / sig1[i] - sig[i]);
        al[i] = 2.0 * s2 / e[i];
        be[i] = -c[i] * 2.0 * s2 / e[i];
        ga[i] = -sig[i] * s2;
        sig[i] = s2;
      }
      CHK(group_hpsi<double>(b, gb, yin, B2, zz, al, be, ga, b->gpart64.p, b->gpart64.cap));
      std::swap(B0, B1);   // X <- Y, Y <- Y1, Y1 <- X
      std::swap(B1, B2);
    }
    if (B1[0] != b->sys[0]->X.p) {   // the result is in B1 for every system alike
      std::vector<const void*> src(B1.begin(), B1.end());
      std::vector<void*> dst(n);
      for (int i = 0; i < n; ++i) dst[i] = b->sys[i]->X.p;
      CHK(group_convert(b, gb, 2, src, dst));
    }
    return PWH_OK;
  }
  if (precision != PWH_FP32)
    return fail(PWH_ERR_UNSUPPORTED, "grouped filter: FP64 and FP32 filters only");
  const bool last64 = degree >= 2 && filter_last_fp64();
  CHK(b->gpart32.ensure(group_part_elems<float>(b)));
  if (last64) CHK(b->gpart64.ensure(group_part_elems<double>(b)));
  std::vector<cx<float>*> B0(n), B1(n), B2(n);
  std::vector<const void*> src(n);
  std::vector<void*> dst(n);
  for (int i = 0; i < n; ++i) {
    B0[i] = b->sys[i]->X32.p;
    B1[i] = b->sys[i]->Y32.p;
    B2[i] = b->sys[i]->Y1_32.p;
    src[i] = b->sys[i]->X.p;
    dst[i] = B0[i];
  }
  CHK(group_convert(b, gb, 0, src, dst));
  const int d32 = last64 ? degree - 1 : degree;
  {
    std::vector<const cx<float>*> in(B0.begin(), B0.end()), Z(n, nullptr);
    for (int i = 0; i < n; ++i) { al[i] = sig1[i] / e[i]; be[i] = -c[i] * sig1[i] / e[i]; ga[i] = 0.0; }
    CHK(group_hpsi<float>(b, gb, in, B1, Z, al, be, ga, b->gpart32.p, b->gpart32.cap));
  }
  for (int id = 2; id <= d32; ++id) {
    std::vector<const cx<float>*> yin(B1.begin(), B1.end()), zz(B0.begin(), B0.end());
    for (int i = 0; i < n; ++i) {
      const double s2 = 1.0 / (2.0 / sig1[i] - sig[i]);
      al[i] = 2.0 * s2 / e[i];
      be[i] = -c[i] * 2.0 * s2 / e[i];
      ga[i] = -sig[i] * s2;
      sig[i] = s2;
    }
    CHK(group_hpsi<float>(b, gb, yin, B2, zz, al, be, ga, b->gpart32.p, b->gpart32.cap));
    std::swap(B0, B1);
    std::swap(B1, B2);
  }
  if (!last64) {
    for (int i = 0; i < n; ++i) { src[i] = B1[i]; dst[i] = b->sys[i]->X.p; }
    return group_convert(b, gb, 1, src, dst);
  }
  // prev = B0, cur = B1 -> A64, B64; FP64 step into X
  for (int i = 0; i < n; ++i) { src[i] = B0[i]; dst[i] = b->sys[i]->A64.p; }
  CHK(group_convert(b, gb, 1, src, dst));
  for (int i = 0; i < n; ++i) { src[i] = B1[i]; dst[i] = b->sys[i]->B64.p; }
  CHK(group_convert(b, gb, 1, src, dst));
  std::vector<const cx<double>*> in(n), Z(n);
  std::vector<cx<double>*> o(n);
  for (int i = 0; i < n; ++i) {
    const double s2 = 1.0 / (2.0 / sig1[i] - sig[i]);
    al[i] = 2.0 * s2 / e[i];
    be[i] = -c[i] * 2.0 * s2 / e[i];
    ga[i] = -sig[i] * s2;
    in[i] = b->sys[i]->B64.p;
    Z[i] = b->sys[i]->A64.p;
    o[i] = b->sys[i]->X.p;
  }
  return group_hpsi<double>(b, gb, in, o, Z, al, be, ga, b->gpart64.p, b->gpart64.cap);
}

int group_filter_issue(pwh_batch* b, const std::vector<GroupLaunch>& L, hipStream_t st) {
  unsigned char* base = b->gdesc.p;
  for (const GroupLaunch& l : L) {
    const void* d = base + l.off;
    switch (l.kind) {
      case 0: HIPCHK((launch_group_convert<double, float>(static_cast<const GroupConv*>(d), l.n, l.blocks, st))); break;
      case 1: HIPCHK((launch_group_convert<float, double>(static_cast<const GroupConv*>(d), l.n, l.blocks, st))); break;
      case 2: HIPCHK((launch_group_convert<double, double>(static_cast<const GroupConv*>(d), l.n, l.blocks, st))); break;
      case 3:
        if (l.fp64) HIPCHK(launch_group_proj<double>(static_cast<const GroupProj<double>*>(d), l.n, l.blocks, l.rblocks, st));
        else HIPCHK(launch_group_proj<float>(static_cast<const GroupProj<float>*>(d), l.n, l.blocks, l.rblocks, st));
        break;
      case 4:
        if (l.fp64) HIPCHK(launch_group_expand<double>(static_cast<const GroupExpand<double>*>(d), l.n, l.blocks, st));
        else HIPCHK(launch_group_expand<float>(static_cast<const GroupExpand<float>*>(d), l.n, l.blocks, st));
        break;
      case 5:
        if (l.fp64)
          HIPCHK(launch_group_fft<double>(l.pass, l.key, static_cast<const GroupHpsi<double>*>(d), l.n, l.blocks,
                                          l.smem, st));
        else
          HIPCHK(launch_group_fft<float>(l.pass, l.key, static_cast<const GroupHpsi<float>*>(d), l.n, l.blocks,
                                         l.smem, st));
        break;
      default:
        return fail(PWH_ERR_ARG, "grouped filter: bad launch kind");
    }
  }
  return PWH_OK;
}
}  // namespace

extern "C" {

int pwh_batch_create(pwh_system* const* systems, const int* nstates, int nsys, pwh_batch** out) {
  if (!systems || !nstates || nsys < 1 || !out) return fail(PWH_ERR_ARG, "bad batch arguments");
  CHK(ensure_init());
  auto b = new pwh_batch();
  for (int i = 0; i < nsys; ++i) {
    if (!systems[i] || nstates[i] < 1) { delete b; return fail(PWH_ERR_ARG, "bad system in batch"); }
    b->sys.push_back(systems[i]);
    b->nb.push_back(nstates[i]);
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      pwh_batch_destroy(b);
      return fail(PWH_ERR_HIP, "batch stream/event creation failed");
    }
    b->streams.push_back(st);
    b->done.push_back(ev);
  }
  HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  *out = b;
  return PWH_OK;
}

int pwh_batch_destroy(pwh_batch* b) {
  if (!b) return PWH_OK;
  if (g.stream) hipStreamSynchronize(g.stream);
  batch_drop_graph(b);
  for (auto st : b->streams) { hipStreamSynchronize(st); hipStreamDestroy(st); }
  for (auto ev : b->done) hipEventDestroy(ev);
  if (b->fork) hipEventDestroy(b->fork);
  delete b;
  return PWH_OK;
}

int pwh_batch_set(pwh_batch* b, int i, const double* X) {
  if (!b || i < 0 || i >= (int)b->sys.size() || !X) return fail(PWH_ERR_ARG, "bad batch index");
  CHK(load_block(b->sys[i], X, b->nb[i]));
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_batch_get(pwh_batch* b, int i, double* X) {
  if (!b || i < 0 || i >= (int)b->sys.size() || !X) return fail(PWH_ERR_ARG, "bad batch index");
  HIPCHK(hipStreamSynchronize(g.stream));
  return store_block(b->sys[i], b->sys[i]->X.p, X, b->nb[i]);
}

int pwh_batch_chebyfilt(pwh_batch* b, int degree, const double* lb, const double* ub, int precision,
                        int mode) {
  if (!b || !lb || !ub || degree < 1) return fail(PWH_ERR_ARG, "bad batch filter arguments");
  if (precision < PWH_FP64 || precision > PWH_BF16X6) return fail(PWH_ERR_ARG, "bad precision");
  const int n = (int)b->sys.size();
  for (int i = 0; i < n; ++i) {
    if (!(ub[i] > lb[i])) return fail(PWH_ERR_ARG, "chebyfilt needs ub > lb");
    CHK(filter_prepare(b->sys[i], b->nb[i], precision));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  if (mode == 0) {
    for (int i = 0; i < n; ++i)
      CHK(filter_on_stream(b->sys[i], b->nb[i], degree, lb[i], ub[i], precision, g.stream));
    return PWH_OK;
  }
  if (mode != 1 && mode != 2) return fail(PWH_ERR_ARG, "bad batch filter mode");
  const bool same = b->exec && b->g_mode == mode && b->g_degree == degree && b->g_prec == precision &&
                    b->g_gen == g.layout_gen && std::equal(lb, lb + n, b->g_lb.begin()) &&
                    std::equal(ub, ub + n, b->g_ub.begin());
  if (!same) {
    batch_drop_graph(b);
    GroupBuilder gb;
    if (mode == 2) {
      CHK(group_filter_build(b, gb, degree, lb, ub, precision));
      CHK(b->gdesc.ensure(gb.host.size()));
      HIPCHK(hipMemcpyAsync(b->gdesc.p, gb.host.data(), gb.host.size(), hipMemcpyHostToDevice, g.stream));
      HIPCHK(hipStreamSynchronize(g.stream));
    }
    const bool timing = g.timing;
    g.timing = false;
    int rc = PWH_OK;
    HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
    if (mode == 2) {
      rc = group_filter_issue(b, gb.launches, g.stream);
    } else {
      hipEventRecord(b->fork, g.stream);
      for (int i = 0; i < n && rc == PWH_OK; ++i) {
        hipStreamWaitEvent(b->streams[i], b->fork, 0);
        rc = filter_on_stream(b->sys[i], b->nb[i], degree, lb[i], ub[i], precision, b->streams[i]);
        hipEventRecord(b->done[i], b->streams[i]);
        hipStreamWaitEvent(g.stream, b->done[i], 0);
      }
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g.stream, &graph);
    g.timing = timing;
    if (rc != PWH_OK) { if (graph) hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(PWH_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(e));
    b->graph = graph;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_degree = degree;
    b->g_prec = precision;
    b->g_mode = mode;
    b->g_gen = g.layout_gen;
    b->g_lb.assign(lb, lb + n);
    b->g_ub.assign(ub, ub + n);
  }
  HIPCHK(hipGraphLaunch(b->exec, g.stream));
  return PWH_OK;
}

int pwh_synchronize(void) {
  CHK(ensure_init());
  HIPCHK(hipStreamSynchronize(g.stream));
  return PWH_OK;
}

int pwh_timing_enable(int on) {
  CHK(ensure_init());
  g.timing = on != 0;
  return PWH_OK;
}

int pwh_timing_reset(void) {
  resolve_timers();
  for (int i = 0; i < TM_COUNT; ++i) { g.total_ms[i] = 0; g.count[i] = 0; }
  return PWH_OK;
}

int pwh_timing_get(const char* name, double* total_ms, int* count) {
  resolve_timers();
  for (int i = 0; i < TM_COUNT; ++i)
    if (name && std::strcmp(name, kTimerNames[i]) == 0) {
      if (total_ms) *total_ms = g.total_ms[i];
      if (count) *count = g.count[i];
      return PWH_OK;
    }
  return fail(PWH_ERR_ARG, "unknown timer name");
}

}  // extern "C"
