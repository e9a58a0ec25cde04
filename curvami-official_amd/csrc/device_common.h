// Device helpers shared by the assembly and energy kernels: the laminate
// stiffness blocks as kernel arguments and the DPP wave reduction.
#pragma once
#include <hip/hip_runtime.h>

namespace cvm {

// A, B, D of the laminate (Constitutive::computeABD), row-major 3 x 3.
struct MatC {
    double A[9], B[9], D[9];
};

// Sum over the 64 lanes of a wave, returned to every lane, with DPP lane moves
// (no LDS traffic): pairs, quads, half rows and rows by quad_perm / mirror
// moves, then rows 0+1 and 2+3 by row_bcast15 and the total by row_bcast31 into
// lane 63, read back as a scalar.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xb1, 0xf>(v);     // quad_perm(1,0,3,2)
    v += dpp_move<0x4e, 0xf>(v);     // quad_perm(2,3,0,1)
    v += dpp_move<0x141, 0xf>(v);    // row_half_mirror
    v += dpp_move<0x140, 0xf>(v);    // row_mirror: every lane holds its row's sum
    v += dpp_move<0x142, 0xa>(v);    // row_bcast15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);    // row_bcast31 into rows 2 and 3: lane 63 holds the total
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

}  // namespace cvm
