// r3d_fma.hip - the FMA-contracted variant of the stage kernel.
//
// The same source as the bit-exact stage kernel (r3d_stage.hpp), compiled in
// this translation unit with -ffp-contract=fast and in namespace r3d_fma, so
// that products feeding sums become fused multiply-adds.  Results differ from
// the reference's rounding; the error against the bit-exact variant is
// measured by tests/test_gpu_fma.py.  Boundary, dt and bookkeeping kernels
// are shared with the bit-exact build.
#include <hip/hip_runtime.h>

#define R3D_KNS r3d_fma
#include "r3d_stage.hpp"
#include "r3d_ctx.hpp"

namespace r3d {

// Launch one fused stage (same arguments and grid as the bit-exact kernel).
hipError_t launch_stage_fma(const StageArgs& a, const Consts& C, const Metrics& M, int ty,
                            bool mag, unsigned grid, hipStream_t stream) {
#define R3D_FMA(TY, S, Z, MG)                                                                     \
    hipLaunchKernelGGL((r3d_fma::k_stage<TY, S, Z, MG, false, false, MG ? TILE_W_MAG : 64>),     \
                       dim3(grid), dim3(r3d_fma::TileShape<TY, MG ? TILE_W_MAG : 64>::NW * 64), 0, stream, a, \
                       C, M)
#define R3D_FMA_STAGES(TY, MG)                                        \
    if (a.stage == 0 && a.zmode == 0) R3D_FMA(TY, 0, 0, MG);          \
    else if (a.stage == 0) R3D_FMA(TY, 0, 1, MG);                     \
    else if (a.stage == 1 && a.zmode == 0) R3D_FMA(TY, 1, 0, MG);     \
    else if (a.stage == 1) R3D_FMA(TY, 1, 2, MG);                     \
    else R3D_FMA(TY, 2, 2, MG);
    if (mag) {
        if (ty != TILE_Y_MAG) return hipErrorInvalidValue;
        R3D_FMA_STAGES(TILE_Y_MAG, true)
    } else {
        if (ty == 8) {
            R3D_FMA_STAGES(8, false)
        } else if (ty == TILE_Y) {
            R3D_FMA_STAGES(TILE_Y, false)
        } else {
            return hipErrorInvalidValue;
        }
    }
#undef R3D_FMA_STAGES
#undef R3D_FMA
    return hipGetLastError();
}

}  // namespace r3d
