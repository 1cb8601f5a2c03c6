// charm_stage_mix_ph0.hip -- stage kernels of the mixed mode: FP64 state,
// FP32 fluxes and integrals, constant-cp Euler (PH = 0) variant
#include "charm_stage.hpp"

namespace charm {

hipError_t launch_stage_mix_ph0(int nc, int flux, const StageArgs &A, const DevParams &P, hipStream_t st)
{
    return launch_stage_t<double, float, 0>(nc, flux, A, P, st);
}

const void *stage_fn_mix_ph0(int nc, int flux, int bs)
{
    return stage_fn_t<double, float, 0>(nc, flux, bs);
}

} // namespace charm
