// pwh_setup_dev.h -- launchers of the per-G set-up kernels in pwh_setup_dev.hip
// (NLFormFactor, FormFactor, PhaseFactor, StructureFactor on the GPU).
#pragma once
#include "pwh_internal.h"

namespace pwh {

hipError_t nl_prep(const double* r, const double* vl, const double* vloc, const double* psi, int m, double* fun,
                   double* fun1, hipStream_t s);
hipError_t nl_wnl(const double* fun1, int m, double h, double fpibo, int l, double* out, hipStream_t s);
// xk: host pointer to the k-point's 3 components (passed by value)
hipError_t nl_pkg(const double* r, const double* fun, int m, double h, int l, int ml, const double* xkg,
                  const double* ggk, const double* xk, int ngw, double tpiba, int nlmax, double* pkg_a,
                  hipStream_t s);
hipError_t ff_prep(const double* r, const double* vion_loc, int m, double zv, double rgauss, double* vloc,
                   hipStream_t s);
hipError_t ff_vps(const double* r, const double* vloc, int m, double h, const double* g2, int ng, double tpiba,
                  double fpibom, double zv, double omega, double facc, int nsp, int is, double* vps, double* rhops,
                  hipStream_t s);
// tables [natoms][3][ldt], ldt >= max(nr) + 2
hipError_t ei_tables(const double* taup, int natoms, const int nr[3], int ldt, cx<double>* ei, hipStream_t s);
hipError_t eigr_dev(const cx<double>* ei, int ldt, const int* na, int nsp, int nax, const int* igk,
                    const int* in123, int ngx, int ngw, int ngwx, const int idx0[3], cx<double>* eigr,
                    hipStream_t s);
hipError_t sfac_dev(const cx<double>* ei, int ldt, const int* na, int nax, int nsp, const int* in123, int ngx,
                    int ng, const int idx0[3], cx<double>* sfac, hipStream_t s);

}  // namespace pwh
