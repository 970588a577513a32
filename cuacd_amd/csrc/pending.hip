// pending.hip - entry points whose kernels are not built yet.
//
// Each returns -1 with an explicit error (never a CPU fallback). Entries move
// out of this file as their kernels land.
#include "ctx.hpp"

extern "C" {

int gpu_kdop_hull(gpu_ctx_t ctx, const float*, int, const int*, int, int, int, float*, int*, int*, int*, float*,
                  int*) {
    cuacd::set_error(ctx, "kdop_hull: not built yet");
    return -1;
}

int gpu_mesh_audit(gpu_ctx_t ctx, const float*, int, const int*, int, unsigned int*) {
    cuacd::set_error(ctx, "mesh_audit: not built yet");
    return -1;
}

int gpu_test_postprocess_dc(gpu_ctx_t ctx, const float*, int, const int*, int, int, int, int, float*, int*, int*,
                            int*, int*) {
    cuacd::set_error(ctx, "postprocess_dc: not built yet");
    return -1;
}

}  // extern "C"
