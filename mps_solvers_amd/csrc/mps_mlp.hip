// mps_mlp.hip -- learned warm start for the pressure Poisson CG (placeholder
// until the MFMA kernels land): reports that no network is loaded.
#include <hip/hip_runtime.h>

#include "../../include/mps_amd.h"
#include "mps_ctx.hpp"

void mps_mlp_free(mps_ctx* c) { (void)c; }
int mps_mlp_guess(mps_ctx* c) { (void)c; return MPS_ERR_STATE; }
