// pwh_dft.h -- the small DFT butterflies (radix 2, 3, 4, 5, 7, 8) and the stage hooks
// shared by the FFT passes with compiled plans (pwh_fft.hip) and with runtime plans
// (pwh_fft_rt.hip).  Conventions of fft_fftw3.cpp:27-38: SIGN = +1 is the inverse
// transform exp(+2 pi i jk / n), unscaled; SIGN = -1 the forward one.
#pragma once
#include "pwh_internal.h"

namespace pwh {

// ---------------------------------------------------------------------------
// Small DFTs, w = exp(SIGN * 2 pi i / R)
// ---------------------------------------------------------------------------
template <typename T, int SIGN>
__device__ __forceinline__ void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
  cx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_si<SIGN>(a1 - a3);
  a0 = t0 + t2; a2 = t0 - t2; a1 = t1 + t3; a3 = t1 - t3;
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cx<T>* v) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
  } else if constexpr (R == 4) {
    dft4<T, SIGN>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T, SIGN>(e0, e1, e2, e3);
    dft4<T, SIGN>(o0, o1, o2, o3);
    const T c = T(0.70710678118654752440084436210485);
    // o1 *= w, o2 *= w^2, o3 *= w^3 with w = (1 + SIGN i)/sqrt2
    cx<T> w1 = mk<T>(c * (o1.x - SIGN * o1.y), c * (o1.y + SIGN * o1.x));
    cx<T> w2 = mul_si<SIGN>(o2);
    cx<T> w3 = mk<T>(c * (-o3.x - SIGN * o3.y), c * (-o3.y + SIGN * o3.x));
    v[0] = e0 + o0; v[4] = e0 - o0;
    v[1] = e1 + w1; v[5] = e1 - w1;
    v[2] = e2 + w2; v[6] = e2 - w2;
    v[3] = e3 + w3; v[7] = e3 - w3;
  } else if constexpr (R == 3) {
    const T h = T(0.86602540378443864676372317075294);
    cx<T> s = v[1] + v[2], d = v[1] - v[2];
    cx<T> t = v[0] - T(0.5) * s;
    cx<T> u = mul_si<SIGN>(h * d);
    v[0] = v[0] + s; v[1] = t + u; v[2] = t - u;
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410229341718282);    // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);   // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);    // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);    // sin(4pi/5)
    cx<T> a0 = v[0];
    cx<T> b1 = v[1] + v[4], b2 = v[2] + v[3], d1 = v[1] - v[4], d2 = v[2] - v[3];
    cx<T> r1 = a0 + c1 * b1 + c2 * b2;
    cx<T> r2 = a0 + c2 * b1 + c1 * b2;
    cx<T> i1 = mul_si<SIGN>(s1 * d1 + s2 * d2);
    cx<T> i2 = mul_si<SIGN>(s2 * d1 - s1 * d2);
    v[0] = a0 + b1 + b2;
    v[1] = r1 + i1; v[4] = r1 - i1;
    v[2] = r2 + i2; v[3] = r2 - i2;
  } else if constexpr (R == 7) {
    // X_m = a0 + sum_k cos(2 pi k m / 7) b_k + SIGN i sum_k sin(2 pi k m / 7) d_k (m = 1..3),
    // X_{7-m} with the sine part negated; b_k = v_k + v_{7-k}, d_k = v_k - v_{7-k}
    const T c1 = T(0.62348980185873353052500488400424);    // cos(2pi/7)
    const T c2 = T(-0.22252093395631440428890256449680);   // cos(4pi/7)
    const T c3 = T(-0.90096886790241912623610231950745);   // cos(6pi/7)
    const T s1 = T(0.78183148246802980870844452667406);    // sin(2pi/7)
    const T s2 = T(0.97492791218182360701813168299393);    // sin(4pi/7)
    const T s3 = T(0.43388373911755812047576833284836);    // sin(6pi/7)
    const cx<T> a0 = v[0];
    const cx<T> b1 = v[1] + v[6], b2 = v[2] + v[5], b3 = v[3] + v[4];
    const cx<T> d1 = v[1] - v[6], d2 = v[2] - v[5], d3 = v[3] - v[4];
    const cx<T> r1 = a0 + c1 * b1 + c2 * b2 + c3 * b3;
    const cx<T> r2 = a0 + c2 * b1 + c3 * b2 + c1 * b3;
    const cx<T> r3 = a0 + c3 * b1 + c1 * b2 + c2 * b3;
    const cx<T> i1 = mul_si<SIGN>(s1 * d1 + s2 * d2 + s3 * d3);
    const cx<T> i2 = mul_si<SIGN>(s2 * d1 - s3 * d2 - s1 * d3);
    const cx<T> i3 = mul_si<SIGN>(s3 * d1 - s1 * d2 + s2 * d3);
    v[0] = a0 + b1 + b2 + b3;
    v[1] = r1 + i1; v[6] = r1 - i1;
    v[2] = r2 + i2; v[5] = r2 - i2;
    v[3] = r3 + i3; v[4] = r3 - i3;
  } else {
    static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8, "unsupported radix");
  }
}

// Stage hooks: which input elements of the first stage are live (the others are known
// zeros and are not read), which outputs of the last stage are stored, and a pointwise
// operation applied to the last stage's outputs in registers.
struct NoHooks {
  __device__ __forceinline__ bool live_in(int, int) const { return true; }
  __device__ __forceinline__ bool live_out(int, int) const { return true; }
  template <typename V>
  __device__ __forceinline__ void post(int, int, V&) const {}
};

// x-direction hooks of the plane pass: inactive i1 columns are zero on input of the
// inverse transform; its outputs are multiplied by V; the forward transform stores only
// the active columns.
template <typename T>
struct XInvHooks {
  const unsigned char* act;  // LDS [N1]
  const T* vrow;             // V of this plane, [N2][N1]
  int n1;
  __device__ __forceinline__ bool live_in(int, int e) const { return act[e] != 0; }
  __device__ __forceinline__ bool live_out(int, int) const { return true; }
  __device__ __forceinline__ void post(int t, int e, cx<T>& v) const {
    const T s = vrow[t * n1 + e];
    v = s * v;
  }
};
struct XFwdHooks {
  const unsigned char* act;
  __device__ __forceinline__ bool live_in(int, int) const { return true; }
  __device__ __forceinline__ bool live_out(int, int e) const { return act[e] != 0; }
  template <typename V>
  __device__ __forceinline__ void post(int, int, V&) const {}
};

struct XLiveHooks {
  const unsigned char* act;
  __device__ __forceinline__ bool live_in(int, int e) const { return act[e] != 0; }
  __device__ __forceinline__ bool live_out(int, int) const { return true; }
  template <typename V>
  __device__ __forceinline__ void post(int, int, V&) const {}
};

}  // namespace pwh
