// charm_stage_f32_ph1.hip -- stage kernels, float arithmetic, Navier-Stokes / real-gas (PH = 1) variant
#include "charm_stage.hpp"

namespace charm {

hipError_t launch_stage_f32_ph1(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    return launch_stage_t<float, float, 1>(nc, flux, A, P, st);
}

const void *stage_fn_f32_ph1(int nc, int flux, int bs)
{
    return stage_fn_t<float, float, 1>(nc, flux, bs);
}

} // namespace charm
