// gp_tables.hpp -- binary128 construction of the GP prediction vectors.
#pragma once

#include "hmhd_internal.hpp"

namespace hmhd {

// GP prediction vectors and quadrature weights in the device table layout:
// pv[kPvSet] holds, for rr = 1..R, [4 faces][nq][ns] at gp_pv_off(rr);
// qw[(kMaxR+1)*4] holds qw[rr*4 + q].  Returns 0 on success, -2 when the
// covariance matrix is not positive definite.
int gp_compute(int R, double ell_over_delta, double xbeg, double xend, double ybeg, double yend, int nx, int ny,
               double* pv, double* qw);

}  // namespace hmhd
