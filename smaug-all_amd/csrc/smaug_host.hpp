// smaug_host.hpp -- host-side helpers shared by the single-grid engine and
// the ensemble: error text, HIP status checks, per-axis metric arrays.
#pragma once
#include "../../include/smaug_engine.h"
#include "smaug_phys.hpp"

#include <hip/hip_runtime.h>
#include <math.h>
#include <vector>

namespace smaug {

// records the message returned by smaug_last_error() and returns code
int fail(int code, const char *fmt, ...);

// validates the fields of one grid's configuration that do not depend on
// other grids (dimensionality, sizes, boundaries, metric pointers)
int check_cfg(const smaug_cfg *cfg);

// along every periodic dim (boundtype 0) the background of each ghost cell
// must equal that of its periodic image: the ghost fill copies only the
// perturbation, so a background that is not periodic would join, say, the
// dense bottom's perturbation to the tenuous top's background (a negative
// full density there).  Returns SMAUG_EINVAL with a message otherwise.
int check_periodic_bg(int D, const int n[3], const int bc[3], const double *bg);

// dx, 1/dx and the separable driver profile along dim d (length n[d]; unit
// arrays along the unused third dim of a 2-D grid).  The driver profile is
// exp(-(z-z0)^2/sz^2) along dim 0 and exp(-(x-xc)^2/sh^2) along the
// horizontal dims, so that the product over dims is the driver's
// Gaussian footprint (usersource.cu:4-92); profq is (x - xc) times the
// profile along the horizontal dims (the torsional driver's lever arms),
// 0 along dim 0.
void axis_arrays(const smaug_cfg *cfg, int d, int nd, std::vector<double> &dx,
                 std::vector<double> &idx, std::vector<double> &prof,
                 std::vector<double> &profq);

// compensated FP64 sums of the evolved fields of u over the interior cells
// of grid g (k_sums plus a compensated host pass), nve values into sums;
// part is device scratch of at least 2 * nve * sums_blocks() doubles
int field_sums(int D, int prec, const Geo &g, const void *u, hipStream_t st, double *part,
               double *sums);
constexpr int sums_blocks() { return 512; }

} // namespace smaug

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return smaug::fail(SMAUG_EHIP, "%s failed: %s (%s:%d)", #x,                   \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)
