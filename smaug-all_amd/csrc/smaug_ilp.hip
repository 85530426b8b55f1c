// smaug_ilp.hip -- the kernels that run faster under LLVM's max-ILP machine
// scheduler, which the Makefile selects for this file alone
// (-amdgpu-sched-strategy=max-ilp): the FP32 z-streamed hyperdiffusive
// sources with the in-plane ratios formed in the kernel (k_hyper_zsr_f) and
// the ratio pairs along k (k_ratio_kz_f, k_ratio_kz_d).  At 256^3 per
// launch: sources 0.811 -> 0.772 ms, FP32 ratios along k 0.263 -> 0.242 ms,
// FP64 0.463 -> 0.443 ms.  The stage update, the FP64 sources and the 2-D
// per-cell update are slower under it (0.855 -> 0.864 ms, 2.00 -> 2.06 ms,
// Orszag-Tang 1024^2 step 0.255 -> 0.283 ms), so they stay in
// smaug_engine.hip under the default scheduler.  The kernels are only
// instantiated here; smaug_engine.hip launches them through these functions.
#define SMAUG_ILP_TU
#include "smaug_stream.hpp"

namespace smaug {

hipError_t launch_hyper_zsr_f(bool bgp, unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                              const Phys &ph, const float *uin, const float *R, const float *aux,
                              const double *bg, const unsigned long long *cmb, float *dUh)
{
    hipLaunchKernelGGL(bgp ? k_hyper_zsr_f<true> : k_hyper_zsr_f<false>, dim3(nb), dim3(256), 0,
                       st, g, zm, ph, uin, R, aux, bg, cmb, dUh);
    return hipGetLastError();
}

hipError_t launch_ratio_kz_f(unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                             const float *u, const float *aux, float *R)
{
    hipLaunchKernelGGL(k_ratio_kz_f, dim3(nb), dim3(KZ_NT), 0, st, g, zm, u, aux, R);
    return hipGetLastError();
}

hipError_t launch_ratio_kz_d(unsigned nb, hipStream_t st, const Geo &g, const ZMap &zm,
                             const double *u, const double *aux, double *R)
{
    hipLaunchKernelGGL(k_ratio_kz_d, dim3(nb), dim3(KZ_NT), 0, st, g, zm, u, aux, R);
    return hipGetLastError();
}

} // namespace smaug
