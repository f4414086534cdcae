// pwh_setup_check.cpp -- a host-only driver of the set-up entry points (pwh_setup.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer (`make -C ffr-pspw-dft-c_amd
// asan`) and run by the CPU test suite.  It covers the index-heavy host code -- the grid
// rule (Setup.cpp:20-34), the half-space enumeration and heapsort (GVectorsGen.cpp,
// sort1.cpp), the G+k maps (GkVectorsGen.cpp) and the phase factors (PhaseFactor.cpp) --
// on cubic, orthorhombic, fcc and hexagonal cells, an odd grid and two k-points, and
// prints one line of integer and floating checksums per cell, which the test compares with
// the same entry points in libpwh.so.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/pwh.h"

struct Cell {
  const char* name;
  double A[9];
  double ecut;
  int nr[3];    // 0: from the grid rule
  int nkpt;
};

int main() {
  const Cell cells[] = {
      {"si8_15ry", {10.266, 0, 0, 0, 10.266, 0, 0, 0, 10.266}, 15.0, {0, 0, 0}, 1},
      {"ortho_22x12x16", {10, 0, 0, 0, 5, 0, 0, 0, 7.2}, 10.0, {0, 0, 0}, 1},
      {"fcc_si2", {5.133, 5.133, 0, 5.133, -5.133, 0, 5.133, 0, 5.133}, 8.0, {0, 0, 0}, 2},
      {"hex_zno", {3.07135, 5.319734248, 0, 6.1427, 0, 0, 0, 0, 9.84291}, 8.0, {0, 0, 0}, 1},
      {"odd_15x21x16", {10.26, 0, 0, 0, 10.26, 0, 0, 0, 10.26}, 6.0, {15, 21, 16}, 1},
  };
  for (const Cell& c : cells) {
    int nr[3] = {c.nr[0], c.nr[1], c.nr[2]};
    if (!nr[0] && pwh_fft_grid(c.A, c.ecut, nr) != PWH_OK) return 1;
    std::vector<double> xk(3 * c.nkpt, 0.0);
    if (c.nkpt > 1) { xk[3] = 0.25; xk[4] = 0.125; xk[5] = -0.375; }
    int ngx = 0, ngwx = 0;
    if (pwh_gvec_sizes(c.A, c.ecut, nr, xk.data(), c.nkpt, &ngx, &ngwx) != PWH_OK) return 2;
    int ng = 0;
    std::vector<int> in123(3 * ngx), ngw(c.nkpt), igk((size_t)ngwx * c.nkpt);
    std::vector<double> xkg((size_t)ngwx * c.nkpt), ggk((size_t)3 * ngwx * c.nkpt);
    std::vector<unsigned> n123((size_t)ngwx * c.nkpt);
    if (pwh_gvectors(c.A, c.ecut, nr, xk.data(), c.nkpt, ngx, ngwx, &ng, in123.data(), ngw.data(), igk.data(),
                     xkg.data(), n123.data(), ggk.data()) != PWH_OK)
      return 3;
    // two species (3 and 2 atoms) at fractional-looking positions in bohr
    const int nsp = 2, nax = 3, na[2] = {3, 2};
    std::vector<double> tau(3 * nax * nsp, 0.0);
    for (int i = 0; i < 3 * nax * nsp; ++i) tau[i] = 0.37 * (i % 7) + 0.11 * (i % 3);
    std::vector<double> eigr((size_t)2 * ngwx * nax * nsp * c.nkpt, 0.0);
    if (pwh_phase_factor(c.A, nr, nsp, na, nax, tau.data(), c.nkpt, ngw.data(), ngwx, igk.data(), ngx,
                         in123.data(), eigr.data()) != PWH_OK)
      return 4;
    long long isum = 0;
    for (int k = 0; k < c.nkpt; ++k)
      for (int i = 0; i < ngw[k]; ++i)
        isum += (long long)(i + 1) * ((long long)n123[(size_t)k * ngwx + i] + 7 * igk[(size_t)k * ngwx + i]);
    double fsum = 0.0, esum = 0.0;
    for (int k = 0; k < c.nkpt; ++k)
      for (int i = 0; i < ngw[k]; ++i) fsum += (i % 13 + 1) * xkg[(size_t)k * ngwx + i];
    for (double v : eigr) esum += std::fabs(v);
    printf("%s %d %d %d %d %d %d %d %lld %.12e %.12e\n", c.name, nr[0], nr[1], nr[2], ng, ngwx, ngw[0],
           c.nkpt > 1 ? ngw[1] : 0, isum, fsum, esum);
  }
  return 0;
}
