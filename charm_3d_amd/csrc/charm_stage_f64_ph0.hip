// charm_stage_f64_ph0.hip -- stage kernels, double arithmetic, constant-cp Euler (PH = 0) variant
#include "charm_stage.hpp"

namespace charm {

hipError_t launch_stage_f64_ph0(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    return launch_stage_t<double, double, 0>(nc, flux, A, P, st);
}

const void *stage_fn_f64_ph0(int nc, int flux, int bs)
{
    return stage_fn_t<double, double, 0>(nc, flux, bs);
}

} // namespace charm
