// pwh_setup.cpp -- host producers of the hot path's constant inputs: the FFT grid, the
// G / G+k index maps and the structure-factor phases.  The integer maps decide which
// grid point every plane wave lands on, so they must equal the reference's bit for
// bit, including the tie order of equal |G|^2 produced by its heapsort.  This file
// is compiled with -ffp-contract=off so every |G|^2 is the same double the reference
// computes (no fused multiply-adds), which is what fixes the sort order.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/pwh.h"

namespace {

struct Lattice {
  double a[3][3];     // a[d] = A{d+1}
  double alat, omega, tpiba, tpiba2, gcut, gcutw;
  double b[3][3];     // reciprocal vectors in 2pi/alat units
};

// Setup.cpp:39-58 and CalcRecipLatt.cpp:5-24.
Lattice make_lattice(const double* A, double ecut) {
  Lattice L;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < 3; ++i) L.a[d][i] = A[3 * d + i];
  const double* a1 = L.a[0];
  const double* a2 = L.a[1];
  const double* a3 = L.a[2];
  L.alat = a1[0];
  L.omega = a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]) - a1[1] * (a2[0] * a3[2] - a2[2] * a3[0]) +
            a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]);
  const double s = L.alat / L.omega;
  L.b[0][0] = (a2[1] * a3[2] - a2[2] * a3[1]) * s;
  L.b[0][1] = (a1[2] * a3[1] - a1[1] * a3[2]) * s;
  L.b[0][2] = (a1[1] * a2[2] - a1[2] * a2[1]) * s;
  L.b[1][0] = (a2[2] * a3[0] - a2[0] * a3[2]) * s;
  L.b[1][1] = (a1[0] * a3[2] - a1[2] * a3[0]) * s;
  L.b[1][2] = (a1[2] * a2[0] - a1[0] * a2[2]) * s;
  L.b[2][0] = (a2[0] * a3[1] - a2[1] * a3[0]) * s;
  L.b[2][1] = (a1[1] * a3[0] - a1[0] * a3[1]) * s;
  L.b[2][2] = (a1[0] * a2[1] - a1[1] * a2[0]) * s;
  L.tpiba = 2. * M_PI / L.alat;
  L.tpiba2 = L.tpiba * L.tpiba;
  L.gcut = 4. * ecut / L.tpiba / L.tpiba;
  L.gcutw = L.gcut / 4.0;
  return L;
}

struct GVec { double g[3]; double gg; };
inline GVec gvec(const Lattice& L, int i, int j, int k) {
  GVec v;
  for (int d = 0; d < 3; ++d) v.g[d] = i * L.b[0][d] + j * L.b[1][d] + k * L.b[2][d];
  v.gg = v.g[0] * v.g[0] + v.g[1] * v.g[1] + v.g[2] * v.g[2];
  return v;
}

// Visits the half space of Miller indices in the reference's order (GVectorsGen.cpp:22-49).
template <typename F>
void half_space(const int* nr, F f) {
  for (int k = 0; k <= nr[2] / 2; ++k) {
    const int j0 = (k == 0) ? 0 : -nr[1] / 2;
    for (int j = j0; j <= nr[1] / 2; ++j) {
      const int i0 = (j == 0 && k == 0) ? 1 : -nr[0] / 2;
      for (int i = i0; i <= nr[0] / 2; ++i) f(i, j, k);
    }
  }
}

// Heapsort index of sort1.cpp:6-55 (FHI98MD order): 1-based ranks of arr[0..n-1] in
// ascending order; not stable, and its tie order is part of the G ordering.
void heap_index(int n, const double* arr, std::vector<int>& index) {
  index.resize(n);
  for (int i = 0; i < n; ++i) index[i] = i + 1;
  if (n < 2) return;
  auto key = [&](int pos) { return arr[index[pos - 1] - 1]; };
  int l = n / 2 + 1, ir = n;
  while (true) {
    int rank;
    double q;
    if (l > 1) {
      --l;
      rank = index[l - 1];
      q = arr[rank - 1];
    } else {
      rank = index[ir - 1];
      q = arr[rank - 1];
      index[ir - 1] = index[0];
      if (--ir == 1) {
        index[0] = rank;
        return;
      }
    }
    int i = l, j = l + l;
    while (j <= ir) {
      if (j < ir && key(j) < key(j + 1)) ++j;
      if (q < key(j)) {
        index[i - 1] = index[j - 1];
        i = j;
        j += j;
      } else {
        j = ir + 1;
      }
    }
    index[i - 1] = rank;
  }
}

// 3x3 inverse with the cofactor order of matinv3x3.cpp (column-major).
void inverse3(const double* hm, double* hi) {
  auto M = [&](int i, int j) { return hm[(j - 1) * 3 + (i - 1)]; };
  double d11 = M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2);
  double d12 = M(2, 3) * M(3, 1) - M(2, 1) * M(3, 3);
  double d13 = M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2);
  double d22 = M(1, 1) * M(3, 3) - M(1, 3) * M(3, 1);
  double d23 = M(1, 1) * M(3, 2) - M(3, 1) * M(1, 2);
  double d33 = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
  double d21 = M(3, 2) * M(1, 3) - M(1, 2) * M(3, 3);
  double d31 = M(1, 2) * M(2, 3) - M(2, 2) * M(1, 3);
  double d32 = M(1, 3) * M(2, 1) - M(1, 1) * M(2, 3);
  double dh = M(1, 1) * d11 + M(1, 2) * d12 + M(1, 3) * d13;
  d23 = -d23;
  auto H = [&](int i, int j) -> double& { return hi[(j - 1) * 3 + (i - 1)]; };
  H(1, 1) = d11 / dh; H(2, 2) = d22 / dh; H(3, 3) = d33 / dh;
  H(1, 2) = d21 / dh; H(1, 3) = d31 / dh; H(2, 3) = d32 / dh;
  H(2, 1) = d12 / dh; H(3, 1) = d13 / dh; H(3, 2) = d23 / dh;
}

}  // namespace

namespace pwh {
// Reduced coordinates tau' = A^-1 tau of n atoms (PhaseFactor.cpp:31-40), the inverse with
// the cofactor order of matinv3x3.cpp; used by the GPU phase and structure factors.
void reduced_coords(const double A[9], const double* tau, int n, double* out) {
  double xmat[9], xmati[9];
  for (int i = 0; i < 9; ++i) xmat[i] = A[i];
  inverse3(xmat, xmati);
  for (int a = 0; a < n; ++a)
    for (int d = 0; d < 3; ++d) {
      const double* t = tau + 3 * a;
      double s1 = xmati[d] * t[0];
      s1 = s1 + xmati[d + 3] * t[1];
      s1 = s1 + xmati[d + 6] * t[2];
      out[3 * a + d] = s1;
    }
}
}  // namespace pwh

extern "C" {

int pwh_fft_grid(const double A[9], double ecut, int nr[3]) {
  if (!A || !nr || !(ecut > 0)) return PWH_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    const double* a = A + 3 * d;
    int n = (int)(2.0 * std::sqrt(ecut * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2])) / M_PI);
    n += 2;
    if (n % 2) ++n;
    nr[d] = n;
  }
  return PWH_OK;
}

int pwh_gvec_sizes(const double A[9], double ecut, const int nr[3], const double* xk, int nkpt,
                   int* ngx, int* ngwx) {
  if (!A || !nr || !xk || nkpt < 1 || !ngx || !ngwx) return PWH_ERR_ARG;
  const Lattice L = make_lattice(A, ecut);
  std::vector<int> ngw(nkpt, 1);
  int jg = 2;
  half_space(nr, [&](int i, int j, int k) {
    const GVec v = gvec(L, i, j, k);
    for (int ik = 0; ik < nkpt; ++ik) {
      const double* x = xk + 3 * ik;
      double dx = v.g[0] - x[0], dy = v.g[1] - x[1], dz = v.g[2] - x[2];
      if (dx * dx + dy * dy + dz * dz <= L.gcutw) ++ngw[ik];
      dx = v.g[0] + x[0]; dy = v.g[1] + x[1]; dz = v.g[2] + x[2];
      if (dx * dx + dy * dy + dz * dz <= L.gcutw) ++ngw[ik];
    }
    if (v.gg < L.gcut) ++jg;
  });
  int n_gx = 2 * (jg - 1) + 1;
  for (int ik = 0; ik < nkpt; ++ik) n_gx = std::max(n_gx, 8 * ngw[ik]);
  int n_gwx = n_gx / 8 + 8;
  *ngx = n_gwx * 8 + 8;
  *ngwx = n_gwx + 1;
  return PWH_OK;
}

int pwh_gvectors(const double A[9], double ecut, const int nr[3], const double* xk, int nkpt,
                 int ngx, int ngwx, int* ng, int* in123, int* ngw, int* igk, double* xkg,
                 unsigned int* n123, double* ggk) {
  if (!A || !nr || !xk || nkpt < 1 || ngx < 1 || ngwx < 1 || !ng || !in123 || !ngw || !igk ||
      !xkg || !n123 || !ggk)
    return PWH_ERR_ARG;
  const Lattice L = make_lattice(A, ecut);
  const int half_cap = ngx / 2 + 1;

  // half-space candidates inside the density sphere, slot 0 reserved for G = 0
  std::vector<double> gsq(1, 0.0);
  std::vector<int> mil(3, 0);
  bool overflow = false;
  half_space(nr, [&](int i, int j, int k) {
    const GVec v = gvec(L, i, j, k);
    if (v.gg <= L.gcut) {
      if ((int)gsq.size() + 1 > half_cap) { overflow = true; return; }
      gsq.push_back(v.gg);
      mil.push_back(i); mil.push_back(j); mil.push_back(k);
    }
  });
  if (overflow) return PWH_ERR_ARG;
  const int nhalf = (int)gsq.size();
  std::vector<int> rank;
  heap_index(nhalf, gsq.data(), rank);
  if (gsq[rank[0] - 1] != 0 || rank[0] != 1) return PWH_ERR_ARG;

  // full list: G = 0, then (+G, -G) pairs in sorted order
  struct Full { int m[3]; double g[3]; int n[3]; };
  std::vector<Full> full;
  full.reserve(2 * nhalf);
  full.push_back(Full{{0, 0, 0}, {0.0, 0.0, 0.0}, {1, 1, 1}});
  for (int s = 1; s < nhalf; ++s) {
    if ((int)full.size() + 2 > ngx) return PWH_ERR_ARG;
    const int id = rank[s] - 1;
    for (int sign = 1; sign >= -1; sign -= 2) {
      Full f;
      for (int d = 0; d < 3; ++d) f.m[d] = sign * mil[3 * id + d];
      if (sign > 0) {
        const GVec v = gvec(L, f.m[0], f.m[1], f.m[2]);
        for (int d = 0; d < 3; ++d) f.g[d] = v.g[d];
      } else {
        for (int d = 0; d < 3; ++d) f.g[d] = -full.back().g[d];
      }
      for (int d = 0; d < 3; ++d) f.n[d] = f.m[d] + 1 + (f.m[d] < 0 ? nr[d] : 0);
      full.push_back(f);
    }
  }
  const int nG = (int)full.size();
  *ng = nG;
  for (int ig = 0; ig < nG; ++ig)
    for (int d = 0; d < 3; ++d) in123[d * ngx + ig] = full[ig].m[d];

  for (int ik = 0; ik < nkpt; ++ik) {
    const double* x = xk + 3 * ik;
    int cnt = 0;
    for (int ig = 0; ig < nG; ++ig) {
      const Full& f = full[ig];
      const double gx = f.g[0] + x[0], gy = f.g[1] + x[1], gz = f.g[2] + x[2];
      const double t = gx * gx + gy * gy + gz * gz;
      if (t <= L.gcutw) {
        if (cnt + 1 > ngwx) return PWH_ERR_ARG;
        const size_t o = (size_t)cnt + (size_t)ngwx * ik;
        igk[o] = ig + 1;
        xkg[o] = t;
        n123[o] = (unsigned int)(f.n[0] + nr[0] * ((f.n[1] - 1) + nr[1] * (f.n[2] - 1)));
        for (int d = 0; d < 3; ++d) ggk[3 * o + d] = f.g[d];
        ++cnt;
      }
    }
    ngw[ik] = cnt;
  }
  return PWH_OK;
}

int pwh_phase_factor(const double A[9], const int nr[3], int nsp, const int* na, int nax,
                     const double* tau, int nkpt, const int* ngw, int ngwx, const int* igk,
                     int ngx, const int* in123, double* eigr_out) {
  if (!A || !nr || nsp < 1 || !na || nax < 1 || !tau || nkpt < 1 || !ngw || !igk || !in123 ||
      !eigr_out)
    return PWH_ERR_ARG;
  typedef std::complex<double> Z;
  Z* eigr = reinterpret_cast<Z*>(eigr_out);
  int half[3], zero[3];
  for (int d = 0; d < 3; ++d) { half[d] = (nr[d] + 1) / 2; zero[d] = half[d]; }  // 0-based slot of n = 0
  double xmat[9], xmati[9];
  for (int i = 0; i < 9; ++i) xmat[i] = A[i];
  inverse3(xmat, xmati);
  // ei[d][(ia + nax*is)*(nr[d]+1) + zero[d] + n] = exp(-2 pi i n tau'_d)
  std::vector<Z> ei[3];
  for (int d = 0; d < 3; ++d) ei[d].assign((size_t)(nr[d] + 1) * nax * nsp, Z(0.0, 0.0));
  for (int is = 0; is < nsp; ++is) {
    for (int ia = 0; ia < na[is]; ++ia) {
      const double* t = tau + 3 * (ia + nax * is);
      for (int d = 0; d < 3; ++d) {
        double s1 = xmati[d] * t[0];
        s1 = s1 + xmati[d + 3] * t[1];
        s1 = s1 + xmati[d + 6] * t[2];
        const double ar = -2.0 * M_PI * s1;
        Z* e = ei[d].data() + (size_t)(nr[d] + 1) * (ia + nax * is) + zero[d];
        e[0] = Z(1.0, 0.0);
        for (int n = 1; n <= half[d]; ++n) {
          const double c = std::cos(n * ar), s = std::sin(n * ar);
          if (zero[d] + n <= nr[d]) e[n] = Z(c, s);   // odd grids: slot never indexed
          e[-n] = Z(c, -s);
        }
      }
    }
  }
  for (int ik = 0; ik < nkpt; ++ik)
    for (int is = 0; is < nsp; ++is)
      for (int ia = 0; ia < nax; ++ia)
        for (int ig = 0; ig < ngw[ik]; ++ig) {
          const int igp = igk[ig + ngwx * ik] - 1;
          const size_t base = (size_t)(ia + nax * is);
          const Z a = ei[0][base * (nr[0] + 1) + zero[0] + in123[igp]];
          const Z b = ei[1][base * (nr[1] + 1) + zero[1] + in123[ngx + igp]];
          const Z c = ei[2][base * (nr[2] + 1) + zero[2] + in123[2 * ngx + igp]];
          // (a*b)*c with the textbook product, as C99 complex multiplication evaluates it
          const double abr = a.real() * b.real() - a.imag() * b.imag();
          const double abi = a.real() * b.imag() + a.imag() * b.real();
          eigr[ig + (size_t)ngwx * (ia + (size_t)nax * (is + (size_t)nsp * ik))] =
              Z(abr * c.real() - abi * c.imag(), abr * c.imag() + abi * c.real());
        }
  return PWH_OK;
}

}  // extern "C"
