// Device helpers shared by the assembly and energy kernels: the laminate
// stiffness blocks as kernel arguments and the wave reductions (DPP moves and
// gfx950 half-wave swaps).
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

// Sixteen wave sums at once, by halving: each exchange step sends half of the
// lane's values to its partner and keeps the other half, so the 16 sums cost
// about 95 instructions instead of 16 separate reductions' 300.
//  - permlane32_swap pairs lanes l and l + 32: afterwards lanes 0-31 hold the
//    sums of v[k] and lanes 32-63 those of v[k + 8] (k < 8);
//  - permlane16_swap pairs rows 0/1 and 2/3 the same way (4 values left);
//  - row_mirror (lane i with 15 - i) and row_half_mirror (i with 7 - i) halve
//    twice more on lane bits 3 and 2, and two quad sums finish;
// value j ends in every lane of the quad at lane 32 h + 16 q + 8 b3 + 4 b2 with
// j = 8 h + 4 q + 2 b3 + b2, from where it is read back as a scalar. All 64
// lanes must be active.
__device__ __forceinline__ double swap32_add(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ double swap16_add(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ void wave_sum16(const double (&v)[16], double (&out)[16]) {
    const int lane = static_cast<int>(__lane_id());
    const bool b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1;
    double w[8], x[4], y[2];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = swap32_add(v[k], v[k + 8]);
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = swap16_add(w[k], w[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double keep = b3 ? x[k + 2] : x[k], send = b3 ? x[k] : x[k + 2];
        y[k] = keep + dpp_move_all<0x140>(send);     // row_mirror
    }
    double z;
    {
        const double keep = b2 ? y[1] : y[0], send = b2 ? y[0] : y[1];
        z = keep + dpp_move_all<0x141>(send);        // row_half_mirror
    }
    z += dpp_move_all<0xb1>(z);                      // quad_perm(1,0,3,2)
    z += dpp_move_all<0x4e>(z);                      // quad_perm(2,3,0,1)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int src = 32 * (j >> 3) + 16 * ((j >> 2) & 1) + 8 * ((j >> 1) & 1) + 4 * (j & 1);
        const int lo = __builtin_amdgcn_readlane(__double2loint(z), src);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(z), src);
        out[j] = __hiloint2double(hi, lo);
    }
}

}  // namespace cvm
