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
// lane 63, read back as a scalar. Every move writes all lanes (a lane whose
// source lies outside the wave reads 0), so no destination value has to be set
// first; the rows that the broadcasts should skip end up with other partial
// sums, but none of them feeds lane 63, whose additions are
// ((r3 + r2) + (r1 + r0)) for the row sums r0..r3.
template <int CTRL>
__device__ __forceinline__ double dpp_move_all(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move_all<0xb1>(v);      // quad_perm(1,0,3,2)
    v += dpp_move_all<0x4e>(v);      // quad_perm(2,3,0,1)
    v += dpp_move_all<0x141>(v);     // row_half_mirror
    v += dpp_move_all<0x140>(v);     // row_mirror: every lane holds its row's sum
    v += dpp_move_all<0x142>(v);     // row_bcast15: row 3 holds r3 + r2, row 1 r1 + r0
    v += dpp_move_all<0x143>(v);     // row_bcast31: lane 63 holds (r3 + r2) + (r1 + r0)
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

}  // namespace cvm
