// pwh_setup_dev.hip -- the per-G set-up work on the GPU: the nonlocal projector form
// factors (NLFormFactor.cpp), the local form factors (FormFactor.cpp), the phase factors
// EIGR (PhaseFactor.cpp) and the structure factor (StructureFactor.cpp).
//
// A radial transform is one wave per G: the lanes stride over the logarithmic mesh, add
// their Simpson-weighted terms, and the wave reduces them in a fixed shuffle tree
// (simpson.cpp's rule h/3 (f_0 + 4 f_odd + 2 f_even + f_last)).  FP64 throughout.
#include "pwh_setup_dev.h"

namespace pwh {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double simpson_w(int i, int n) {
  if (i == 0 || i == n - 1) return 1.0;
  return (i & 1) ? 4.0 : 2.0;
}

// j_l-type factors of NLFormFactor.cpp:62-92 (l = 1: sin x / x; 2: (sin x / x - cos x) / x;
// 3: ((3 / x^2 - 1) sin x - 3 / x cos x) / x)
__device__ __forceinline__ double bessel_fac(int l, double x) {
  double s, c;
  sincos(x, &s, &c);
  if (l == 1) return s / x;
  if (l == 2) return (s / x - c) / x;
  return ((3.0 / (x * x) - 1.0) * s - 3.0 / x * c) / x;
}

}  // namespace

// fun[ir] = (VION_l - VION_loc) PSI_l r^2 and fun1[ir] = (VION_l - VION_loc) PSI_l^2 r
// (NLFormFactor.cpp:36-42) for one species and channel
__global__ void k_nl_prep(const double* __restrict__ r, const double* __restrict__ vl,
                          const double* __restrict__ vloc, const double* __restrict__ psi, int m,
                          double* __restrict__ fun, double* __restrict__ fun1) {
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double rr = r[i];
    const double dv = vl[i] - vloc[i];
    fun[i] = dv * psi[i] * (rr * rr);
    fun1[i] = dv * (psi[i] * psi[i]) * rr;
  }
}

// WNL for the channel: (2l - 1) 4 pi / Omega / simpson(fun1) (NLFormFactor.cpp:43-50)
__global__ void k_nl_wnl(const double* __restrict__ fun1, int m, double h, double fpibo, int l, double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += 64) s += simpson_w(i, m) * fun1[i];
  s = wave_sum(s);
  if (threadIdx.x == 0) *out = (2.0 * l - 1.0) * fpibo / (s * h / 3.0);
}

// One wave per G of k-point ik: the channel's 2l - 1 components of PKG_A(m, ig, is)
// (NLFormFactor.cpp:52-124)
__global__ __launch_bounds__(256) void k_nl_pkg(const double* __restrict__ r, const double* __restrict__ fun, int m,
                                                double h, int l, int ml, const double* __restrict__ xkg,
                                                const double* __restrict__ ggk, double xk0, double xk1, double xk2,
                                                int ngw, double tpiba, int nlmax, double* __restrict__ pkg_a) {
  const int ig = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ig >= ngw) return;
  const double t = sqrt(xkg[ig]);
  const double arg = t * tpiba;
  const bool small = t < 1.0e-4;
  double* out = pkg_a + (size_t)nlmax * ig + ml;
  if (small && l > 1) {
    if (lane < 2 * l - 1) out[lane] = 0.0;
    return;
  }
  double s = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double f = small ? fun[i] : bessel_fac(l, arg * r[i]) * fun[i];
    s += simpson_w(i, m) * f;
  }
  s = wave_sum(s);
  const double q1 = s * h / 3.0;
  if (lane != 0) return;
  if (l == 1) {
    out[0] = q1;
    return;
  }
  const double cx = (ggk[3 * ig + 0] + xk0) / t, cy = (ggk[3 * ig + 1] + xk1) / t, cz = (ggk[3 * ig + 2] + xk2) / t;
  if (l == 2) {
    out[0] = q1 * cx;
    out[1] = q1 * cy;
    out[2] = q1 * cz;
  } else {
    const double sq3 = 1.7320508075688772;
    out[0] = q1 * 0.5 * (3.0 * (cz * cz) - 1.0);
    out[1] = q1 * sq3 * cz * cx;
    out[2] = q1 * sq3 * cz * cy;
    out[3] = q1 * sq3 * cx * cy;
    out[4] = q1 * sq3 / 2.0 * ((cx * cx) - (cy * cy));
  }
}

// vloc[ir] = VION_loc + ZV / r erf(r / rgauss) (FormFactor.cpp:42-46)
__global__ void k_ff_prep(const double* __restrict__ r, const double* __restrict__ vion_loc, int m, double zv,
                          double rgauss, double* __restrict__ vloc) {
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double vv = -zv / r[i] * erf(r[i] / rgauss);
    vloc[i] = vion_loc[i] - vv;
  }
}

// One wave per density G: VPS(is, ig) = 4 pi / Omega simpson(vloc sin(arg)/arg r^3),
// arg = r |G| tpiba (G = 0: without the sinc); RHOPS(is, ig) = -ZV/Omega exp(-facc G) for
// ig >= 1 (FormFactor.cpp:27-71)
__global__ __launch_bounds__(256) void k_ff_vps(const double* __restrict__ r, const double* __restrict__ vloc, int m,
                                                double h, const double* __restrict__ g2, int ng, double tpiba,
                                                double fpibom, double zv, double omega, double facc, int nsp,
                                                int is, double* __restrict__ vps, double* __restrict__ rhops) {
  const int ig = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ig >= ng) return;
  const double g = sqrt(g2[ig]) * tpiba;
  double s = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double rr = r[i];
    const double r3 = rr * rr * rr;
    double f;
    if (ig == 0) {
      f = vloc[i] * r3;
    } else {
      const double a = rr * g;
      f = vloc[i] * sin(a) / a * r3;
    }
    s += simpson_w(i, m) * f;
  }
  s = wave_sum(s);
  if (lane != 0) return;
  vps[is + (size_t)nsp * ig] = (s * h / 3.0) * fpibom;
  if (ig > 0) rhops[is + (size_t)nsp * ig] = -zv / omega * exp(-facc * g2[ig]);
}

// PhaseFactor.cpp:44-61: EI_d(idx0 +- i) = cos(i ar_d) +- i sin(i ar_d), EI_d(idx0) = 1, for
// the three directions of one atom; tables [nsp*nax][3][ldt]
__global__ void k_ei_tables(const double* __restrict__ taup, int nr1, int nr2, int nr3, int ldt,
                            cx<double>* __restrict__ ei) {
  const int atom = blockIdx.x;
  const int d = blockIdx.y;
  const int nr = d == 0 ? nr1 : (d == 1 ? nr2 : nr3);
  const int idx0 = (nr + 1) / 2;   // 0-based position of index zero
  const double ar = -2.0 * 3.14159265358979323846 * taup[3 * atom + d];
  cx<double>* t = ei + ((size_t)atom * 3 + d) * ldt;
  for (int i = threadIdx.x; i <= (nr + 1) / 2; i += blockDim.x) {
    if (i == 0) {
      t[idx0] = mk<double>(1.0, 0.0);
      continue;
    }
    double s, c;
    sincos(i * ar, &s, &c);
    t[idx0 + i] = mk<double>(c, s);
    t[idx0 - i] = mk<double>(c, -s);
  }
}

__device__ __forceinline__ cx<double> ei_prod(const cx<double>* t, int ldt, int n1, int n2, int n3) {
  const cx<double> a = t[n1], b = t[ldt + n2], c = t[2 * ldt + n3];
  const cx<double> ab = mk<double>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  return mk<double>(ab.x * c.x - ab.y * c.y, ab.x * c.y + ab.y * c.x);
}

// EIGR(ig, ia, is, ik) = EI1 EI2 EI3 at the G+k vector's IN1..IN3 (PhaseFactor.cpp:64-77);
// atoms past NA(is) give 0 (their tables are never filled in the reference)
__global__ void k_eigr(const cx<double>* __restrict__ ei, int ldt, const int* __restrict__ na, int nax,
                       const int* __restrict__ igk, const int* __restrict__ in123, int ngx, int ngw,
                       int ngwx, int idx01, int idx02, int idx03, cx<double>* __restrict__ eigr) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  const int ia = blockIdx.y, is = blockIdx.z;
  if (ig >= ngw) return;
  cx<double> v = mk<double>(0.0, 0.0);
  if (ia < na[is]) {
    const int g = igk[ig] - 1;
    v = ei_prod(ei + ((size_t)(is * nax + ia) * 3) * ldt, ldt, idx01 + in123[g], idx02 + in123[ngx + g],
                idx03 + in123[2 * ngx + g]);
  }
  eigr[ig + (size_t)ngwx * (ia + (size_t)nax * is)] = v;
}

// SFAC(is, ig) = sum over the species' atoms of EI1 EI2 EI3 (StructureFactor.cpp:18-27)
__global__ void k_sfac(const cx<double>* __restrict__ ei, int ldt, const int* __restrict__ na, int nax, int nsp,
                       const int* __restrict__ in123, int ngx, int ng, int idx01, int idx02, int idx03,
                       cx<double>* __restrict__ sfac) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  const int is = blockIdx.y;
  if (ig >= ng) return;
  cx<double> s = mk<double>(0.0, 0.0);
  for (int ia = 0; ia < na[is]; ++ia) {
    const cx<double> v = ei_prod(ei + ((size_t)(is * nax + ia) * 3) * ldt, ldt, idx01 + in123[ig],
                                 idx02 + in123[ngx + ig], idx03 + in123[2 * ngx + ig]);
    s = mk<double>(s.x + v.x, s.y + v.y);
  }
  sfac[is + (size_t)nsp * ig] = s;
}

// ---------------------------------------------------------------------------
hipError_t nl_prep(const double* r, const double* vl, const double* vloc, const double* psi, int m, double* fun,
                   double* fun1, hipStream_t s) {
  hipLaunchKernelGGL(k_nl_prep, dim3(1), dim3(256), 0, s, r, vl, vloc, psi, m, fun, fun1);
  return hipGetLastError();
}
hipError_t nl_wnl(const double* fun1, int m, double h, double fpibo, int l, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_nl_wnl, dim3(1), dim3(64), 0, s, fun1, m, h, fpibo, l, out);
  return hipGetLastError();
}
hipError_t nl_pkg(const double* r, const double* fun, int m, double h, int l, int ml, const double* xkg,
                  const double* ggk, const double* xk, int ngw, double tpiba, int nlmax, double* pkg_a,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_nl_pkg, dim3((ngw + 3) / 4), dim3(256), 0, s, r, fun, m, h, l, ml, xkg, ggk, xk[0], xk[1],
                     xk[2], ngw, tpiba, nlmax, pkg_a);
  return hipGetLastError();
}
hipError_t ff_prep(const double* r, const double* vion_loc, int m, double zv, double rgauss, double* vloc,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_ff_prep, dim3(1), dim3(256), 0, s, r, vion_loc, m, zv, rgauss, vloc);
  return hipGetLastError();
}
hipError_t ff_vps(const double* r, const double* vloc, int m, double h, const double* g2, int ng, double tpiba,
                  double fpibom, double zv, double omega, double facc, int nsp, int is, double* vps, double* rhops,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_ff_vps, dim3((ng + 3) / 4), dim3(256), 0, s, r, vloc, m, h, g2, ng, tpiba, fpibom, zv, omega,
                     facc, nsp, is, vps, rhops);
  return hipGetLastError();
}
hipError_t ei_tables(const double* taup, int natoms, const int nr[3], int ldt, cx<double>* ei, hipStream_t s) {
  hipLaunchKernelGGL(k_ei_tables, dim3(natoms, 3), dim3(64), 0, s, taup, nr[0], nr[1], nr[2], ldt, ei);
  return hipGetLastError();
}
hipError_t eigr_dev(const cx<double>* ei, int ldt, const int* na, int nsp, int nax, const int* igk,
                    const int* in123, int ngx, int ngw, int ngwx, const int idx0[3], cx<double>* eigr,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_eigr, dim3((ngw + 255) / 256, nax, nsp), dim3(256), 0, s, ei, ldt, na, nax, igk, in123, ngx,
                     ngw, ngwx, idx0[0], idx0[1], idx0[2], eigr);
  return hipGetLastError();
}
hipError_t sfac_dev(const cx<double>* ei, int ldt, const int* na, int nax, int nsp, const int* in123, int ngx,
                    int ng, const int idx0[3], cx<double>* sfac, hipStream_t s) {
  hipLaunchKernelGGL(k_sfac, dim3((ng + 255) / 256, nsp), dim3(256), 0, s, ei, ldt, na, nax, nsp, in123, ngx, ng,
                     idx0[0], idx0[1], idx0[2], sfac);
  return hipGetLastError();
}

}  // namespace pwh
