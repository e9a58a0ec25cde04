// Kirchhoff-Love residual and block-upper tangent on gfx950 (FP64), with a
// deterministic scatter into the CSC of K.
//
// Reference: NonLinearAssembler::updateStiffness (plate/Assembly.h:158-345),
// Kinematics::compute (plate/Kinematics.h:61-120), Constitutive::evaluate
// (plate/Constitutive.h:57-70), Constraints::evaluateElasticFold
// (plate/Constraints.h:341-436), Meshless::assembleSystem (plate/Meshless.h:113-127).
//
// Element block, regrouped. With Bm_i = Psi_i F0^T (Psi_i = [[px,0],[0,py],[py,px]]),
// Bb_i = beta_i b^T and Db_i = -g_i b^T (g_i = F0 C0^-1 grad Psi_i), the pair block
// of Assembly.h:290-300 equals
//   K_IJ = W [ F0 (Psi_i^T A Psi_j) F0^T + x b^T + b y^T + z b b^T + geo I ],
//   x = F0 Psi_i^T (B beta_j) - Tu_i g_j,  y = F0 Psi_j^T (B beta_i) - Tu_j g_i,
//   z = beta_i^T D beta_j + (b.v) g_i . g_j,  geo = gradPsi_i^T N gradPsi_j,
// (A, B, D symmetric; g_i . g_j = gradPsi_i^T C0^-1 gradPsi_j since C0 = F0^T F0),
// which needs 27 numbers per neighbour and ~110 flop per pair instead of the 3x3
// products of the reference.
//
// Scatter: a workgroup owns a tile of consecutive face GPs. Its GPs are taken
// in order; the pairs of one GP hit distinct slots, so they are added to the
// tile's LDS accumulator in parallel without atomics, then the tile's touched
// blocks go to a scratch row. k_reduce_K sums, per K block, its scratch rows in
// ascending tile order and then its fold contributions in crease-GP order.
#include <algorithm>
#include <initializer_list>
#include <mutex>
#include <string>

#include "cvm_internal.h"
#include "device_common.h"
#include "fold_angle.h"

namespace cvm {

constexpr int TILE_THREADS = 512;    // default workgroup, 8 waves: wave w prepares GPs w and w + 8 of the tile
constexpr int TILE_GPS = 16;         // GPs per tile (presence masks are 16-bit)
constexpr int UMAX = 256;            // union nodes per tile (uint8 union indices, at most 255 used)
constexpr int GS = 17;               // GP record fields (tangent type)
// GP record stride, GP-major: the 16 GPs of a tile start on distinct LDS banks
// (FP64: 17 doubles = 34 dwords apart; FP32: 18 floats, even so that field
// pairs stay 8-byte aligned for paired loads), so lanes on different GPs read
// their GP's fields without bank conflicts
template <typename TT> constexpr int gp_stride() { return sizeof(TT) == 8 ? 17 : 18; }
constexpr uint16_t NO_REC = 0xffff;  // loc entry of a union node that a GP does not see
// Neighbour records (tangent-type elements). With up to 64 neighbours per GP
// (GREC) a record also carries g = F0 C0^-1 grad Psi: 9 fields, FP64 records of
// 18 dwords, FP32 records padded to 10 floats (40 bytes) so that they stay
// 8-byte aligned for paired loads. Wider supports keep the compact record (6
// fields; FP64 padded to 7, FP32 at 6) and rebuild g per pair from the GP's
// F0 and C0^-1, so that a tile of 16 GPs with ~110 neighbours each still fits
// the LDS. All four layouts start records on 32 different LDS bank positions.
template <typename TT, bool GREC> constexpr int rec_stride() {
    return GREC ? (sizeof(TT) == 8 ? 9 : 10) : (sizeof(TT) == 8 ? 7 : 6);
}

// neighbour record planes: grad Psi, beta (bending B-row factor,
// Assembly.h:240-247), W Tu (:265-267) and (GREC) g = F0 C0^-1 grad Psi in the
// tangent type TT, and the residual contribution f_local (:249) in FP64
enum { R_PX = 0, R_PY, R_BE0, R_BE1, R_BE2, R_TU, R_G0, R_G1, R_G2 };
enum { R_F0 = 0, R_F1, R_F2 };
// GP record fields: weight, F0 columns, normal, b.v_curv, W N (membrane resultant times
// the weight), C0^-1
enum { G_W = 0, G_A1 = 1, G_A2 = 4, G_B = 7, G_BDV = 10, G_WN = 11, G_CI = 14 };

// Per tile: (1) wave w prepares GP w of the tile (kinematics, resultants and the
// per-neighbour records, Assembly.h:210-268); (2) every thread owns touched
// blocks (lo, hi) of the tile and sums the GPs that couple them, in ascending GP
// order, from the records of the two nodes: no LDS accumulation, no atomics on
// values; (3) the tile's residual per union node, again in GP order.
// The records hold grad Psi, beta, W Tu and g; D beta (and B beta when B != 0)
// are recomputed per pair from the laminate, which keeps the records (and the
// LDS traffic per pair) small.
// TT = double: the FP64 tangent; TT = float: the mixed-precision tangent (FP32
// pair blocks and FP32 accumulation), with kinematics, resultants and the
// residual still in FP64.
// Workgroups are dispatched round-robin over the 8 XCDs (b -> XCD b % 8). Map
// them so that each XCD runs a contiguous range of tiles in order: neighbouring
// tiles write neighbouring scratch rows of the same K blocks, and those partial
// cache lines then merge in one XCD's L2 instead of being written back from
// several.
__device__ __forceinline__ int xcd_tile(int b, int nwg) {
    constexpr int NX = 8;
    const int q = nwg / NX, r = nwg % NX, x = b % NX;
    return x * q + min(x, r) + b / NX;
}

// Laminate classes (template argument LAM): LAM_ORTHO, B = 0 and A, D without
// the 16/26 couplings (isotropic plies, cross-ply stacks at 0 and 90 degrees
// whose 16/26 terms vanish); LAM_SYM, B = 0 with full A and D; LAM_COUPLED,
// B != 0. The orthotropic class skips the products with those zero entries.
enum { LAM_ORTHO = 0, LAM_SYM = 1, LAM_COUPLED = 2 };

// One tile (the workgroup's NT threads; LDS at smem).
template <int LAM, typename TT, bool WIDE, int NT>
__device__ __forceinline__ void tile_body(
    const int* __restrict__ tile_gp, const int* __restrict__ tile_uoff, const int* __restrict__ tile_unodes,
    const int64_t* __restrict__ tile_toff, const uint16_t* __restrict__ tpairs, const int* __restrict__ tile_nd,
    const uint32_t* __restrict__ slot_dst,
    const int* __restrict__ off, const uint8_t* __restrict__ uloc, const uint8_t* __restrict__ urank,
    const double* __restrict__ sf, int64_t npairs, const double* __restrict__ w,
    const double* __restrict__ xu, const double* __restrict__ kbar, MatC mat, int rec_cap, int umax,
    int with_tangent, TT* __restrict__ scr, double* __restrict__ fscr, const int t, double* smem) {
    const int tid = threadIdx.x;
    const int g0 = tile_gp[t], ngp = tile_gp[t + 1] - g0;
    const int ub = tile_uoff[t], U = tile_uoff[t + 1] - ub;
    const int64_t tb = tile_toff[t];
    const int T = static_cast<int>(tile_toff[t + 1] - tb);
    const int pb = off[g0];

    double* recf = smem;                                        // 3 FP64 planes of rec_cap
    double* uc = recf + static_cast<size_t>(rec_cap) * 3;       // umax * 3
    TT* rect = reinterpret_cast<TT*>(uc + umax * 3);            // rec_cap records of RS TT
    constexpr bool GREC = !WIDE;
    constexpr int RS = rec_stride<TT, GREC>();
    constexpr int GSP = gp_stride<TT>();
    TT* gs = rect + static_cast<size_t>(rec_cap) * RS;          // TILE_GPS records of GSP
    uint32_t* pm = reinterpret_cast<uint32_t*>(gs + TILE_GPS * GSP);  // umax presence masks
    uint16_t* loc = reinterpret_cast<uint16_t*>(pm + umax);     // ngp rows of U record indices
    int* chunk_next = reinterpret_cast<int*>(loc + TILE_GPS * umax);  // next unclaimed item chunk
    auto RF = [&](int f, int r) -> double& { return recf[static_cast<size_t>(f) * rec_cap + r]; };
    auto R = [&](int f, int r) -> TT& { return rect[static_cast<size_t>(r) * RS + f]; };  // record-major
    auto Gp = [&](int f, int wv) -> TT& { return gs[wv * GSP + f]; };

    if (tid < U) {
        const int v = tile_unodes[ub + tid];
        uc[3 * tid] = xu[3 * v]; uc[3 * tid + 1] = xu[3 * v + 1]; uc[3 * tid + 2] = xu[3 * v + 2];
        pm[tid] = 0u;
    }
    if (tid == 0) *chunk_next = NT / 64;
    for (int k = tid; k < ngp * U; k += NT) loc[k] = NO_REC;
    __syncthreads();

    const double* A = mat.A;
    const double* B = mat.B;
    const double* D = mat.D;
    for (int wv = tid >> 6; wv < ngp; wv += NT / 64) {
        const int lane = tid & 63;
        {
            const int g = g0 + wv;
            const int b = off[g], n = off[g + 1] - b;
            const int rb = b - pb;
            const bool act = lane < n;
            double px = 0, py = 0, hxx = 0, hyy = 0, hxy = 0;
            int ul = 0, rk = 0;
            double u0 = 0, u1 = 0, u2 = 0;
            if (act) {
                const int64_t p = b + lane;
                px = sf[npairs + p]; py = sf[2 * npairs + p];
                hxx = sf[3 * npairs + p]; hyy = sf[4 * npairs + p]; hxy = sf[5 * npairs + p];
                ul = uloc[p]; rk = urank[p];
                u0 = uc[3 * ul]; u1 = uc[3 * ul + 1]; u2 = uc[3 * ul + 2];
            }
            // ---- kinematics (Kinematics.h:75-113) ------------------------------------
            // lane k holds neighbour k; neighbours 64.. (large supports) are folded
            // into the lanes' partial sums before the wave reductions
            double p_a1x = u0 * px, p_a1y = u1 * px, p_a1z = u2 * px;
            double p_a2x = u0 * py, p_a2y = u1 * py, p_a2z = u2 * py;
            double p_xx0 = u0 * hxx, p_xx1 = u1 * hxx, p_xx2 = u2 * hxx;
            double p_yy0 = u0 * hyy, p_yy1 = u1 * hyy, p_yy2 = u2 * hyy;
            double p_xy0 = u0 * hxy, p_xy1 = u1 * hxy, p_xy2 = u2 * hxy;
            if constexpr (WIDE) for (int k = lane + 64; k < n; k += 64) {
                const int64_t p = b + k;
                const double qx = sf[npairs + p], qy = sf[2 * npairs + p];
                const double qxx = sf[3 * npairs + p], qyy = sf[4 * npairs + p], qxy = sf[5 * npairs + p];
                const int ql = uloc[p];
                const double v0 = uc[3 * ql], v1 = uc[3 * ql + 1], v2 = uc[3 * ql + 2];
                p_a1x += v0 * qx; p_a1y += v1 * qx; p_a1z += v2 * qx;
                p_a2x += v0 * qy; p_a2y += v1 * qy; p_a2z += v2 * qy;
                p_xx0 += v0 * qxx; p_xx1 += v1 * qxx; p_xx2 += v2 * qxx;
                p_yy0 += v0 * qyy; p_yy1 += v1 * qyy; p_yy2 += v2 * qyy;
                p_xy0 += v0 * qxy; p_xy1 += v1 * qxy; p_xy2 += v2 * qxy;
            }
            const double pv[16] = {p_a1x, p_a1y, p_a1z, p_a2x, p_a2y, p_a2z, p_xx0, p_xx1,
                                   p_xx2, p_yy0, p_yy1, p_yy2, p_xy0, p_xy1, p_xy2, 0.0};
            double sv[16];
            wave_sum16(pv, sv);
            const double a1x = 1.0 + sv[0], a1y = sv[1], a1z = sv[2];
            const double a2x = sv[3], a2y = 1.0 + sv[4], a2z = sv[5];
            const double rxx0 = sv[6], rxx1 = sv[7], rxx2 = sv[8];
            const double ryy0 = sv[9], ryy1 = sv[10], ryy2 = sv[11];
            const double rxy0 = sv[12], rxy1 = sv[13], rxy2 = sv[14];
            const double C00 = a1x * a1x + a1y * a1y + a1z * a1z;
            const double C11 = a2x * a2x + a2y * a2y + a2z * a2z;
            const double C01 = a1x * a2x + a1y * a2y + a1z * a2z;
            const double idet = 1.0 / (C00 * C11 - C01 * C01);
            const double Ci00 = C11 * idet, Ci01 = -C01 * idet, Ci11 = C00 * idet;
            const double nx = a1y * a2z - a1z * a2y, ny = a1z * a2x - a1x * a2z, nz = a1x * a2y - a1y * a2x;
            const double J = sqrt(nx * nx + ny * ny + nz * nz);
            const double bx = nx / J, by = ny / J, bz = nz / J;
            // strains (Voigt, engineering shear) and resultants (Constitutive.h:57-70)
            // on the elastic bending strain E1_el = E1 + kappa_bar (Kinematics.h:114-115)
            const double e0[3] = {0.5 * (C00 - 1.0), 0.5 * (C11 - 1.0), C01};
            double e1[3] = {-(rxx0 * bx + rxx1 * by + rxx2 * bz), -(ryy0 * bx + ryy1 * by + ryy2 * bz),
                            -2.0 * (rxy0 * bx + rxy1 * by + rxy2 * bz)};
            if (kbar) e1[0] += kbar[g];
            double N[3], M[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                N[r] = A[3 * r] * e0[0] + A[3 * r + 1] * e0[1] + A[3 * r + 2] * e0[2];
                M[r] = D[3 * r] * e1[0] + D[3 * r + 1] * e1[1] + D[3 * r + 2] * e1[2];
                if constexpr (LAM == LAM_COUPLED) {
                    N[r] += B[3 * r] * e1[0] + B[3 * r + 1] * e1[1] + B[3 * r + 2] * e1[2];
                    M[r] += B[r] * e0[0] + B[3 + r] * e0[1] + B[6 + r] * e0[2];
                }
            }
            const double W = w[g];
            // V = F0 C0^-1 (Assembly.h:217), v_curv, p_XY, b.v (:221-224)
            const double V0x = a1x * Ci00 + a2x * Ci01, V0y = a1y * Ci00 + a2y * Ci01, V0z = a1z * Ci00 + a2z * Ci01;
            const double V1x = a1x * Ci01 + a2x * Ci11, V1y = a1y * Ci01 + a2y * Ci11, V1z = a1z * Ci01 + a2z * Ci11;
            const double vcx = rxx0 * M[0] + 2.0 * rxy0 * M[2] + ryy0 * M[1];
            const double vcy = rxx1 * M[0] + 2.0 * rxy1 * M[2] + ryy1 * M[1];
            const double vcz = rxx2 * M[0] + 2.0 * rxy2 * M[2] + ryy2 * M[1];
            const double ft0 = a1x * vcx + a1y * vcy + a1z * vcz, ft1 = a2x * vcx + a2y * vcy + a2z * vcz;
            const double pX = Ci00 * ft0 + Ci01 * ft1, pY = Ci01 * ft0 + Ci11 * ft1;
            const double bdv = bx * vcx + by * vcy + bz * vcz;
            // per-neighbour quantities (Assembly.h:233-268)
            // the GP's records in union order (rank rk of the node among the GP's
            // neighbours): a wave's slots, consecutive union nodes, then read
            // consecutive records
            auto record = [&](int rk, double px, double py, double hxx, double hyy, double hxy, int ul) {
                const double gI0 = V0x * px + V1x * py, gI1 = V0y * px + V1y * py, gI2 = V0z * px + V1z * py;
                const double SXX = rxx0 * gI0 + rxx1 * gI1 + rxx2 * gI2;
                const double SYY = ryy0 * gI0 + ryy1 * gI1 + ryy2 * gI2;
                const double SXY = rxy0 * gI0 + rxy1 * gI1 + rxy2 * gI2;
                const double be0 = -hxx + SXX, be1 = -hyy + SYY, be2 = 2.0 * (-hxy + SXY);
                const int r = rb + rk;
                // f_local = (Bm^T N + Bb^T M) W (:249-250)
                const double ca = px * N[0] + py * N[2], cb = py * N[1] + px * N[2];
                const double cm = be0 * M[0] + be1 * M[1] + be2 * M[2];
                RF(R_F0, r) = W * (a1x * ca + a2x * cb + bx * cm);
                RF(R_F1, r) = W * (a1y * ca + a2y * cb + by * cm);
                RF(R_F2, r) = W * (a1z * ca + a2z * cb + bz * cm);
                R(R_PX, r) = static_cast<TT>(px); R(R_PY, r) = static_cast<TT>(py);
                R(R_BE0, r) = static_cast<TT>(be0); R(R_BE1, r) = static_cast<TT>(be1);
                R(R_BE2, r) = static_cast<TT>(be2);
                R(R_TU, r) = static_cast<TT>(W * ((px * pX + py * pY) - (hxx * M[0] + 2.0 * hxy * M[2] + hyy * M[1])));
                if constexpr (GREC) {
                    R(R_G0, r) = static_cast<TT>(gI0); R(R_G1, r) = static_cast<TT>(gI1);
                    R(R_G2, r) = static_cast<TT>(gI2);
                }
                loc[wv * U + ul] = static_cast<uint16_t>(r);   // the record, tile-relative
                atomicOr(&pm[ul], 1u << wv);
            };
            if (act) record(rk, px, py, hxx, hyy, hxy, ul);
            if constexpr (WIDE) for (int k = lane + 64; k < n; k += 64) {
                const int64_t p = b + k;
                record(urank[p], sf[npairs + p], sf[2 * npairs + p], sf[3 * npairs + p], sf[4 * npairs + p],
                       sf[5 * npairs + p], uloc[p]);
            }
            if (lane == 0) {
                const double gv[GS] = {W, a1x, a1y, a1z, a2x, a2y, a2z, bx, by, bz, bdv, W * N[0], W * N[1],
                                       W * N[2], Ci00, Ci01, Ci11};
#pragma unroll
                for (int f = 0; f < GS; ++f) Gp(f, wv) = static_cast<TT>(gv[f]);
            }
        }
    }
    __syncthreads();

    if (with_tangent) {
        TT Ac[9], Bc[9], Dc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Ac[k] = static_cast<TT>(A[k]); Bc[k] = static_cast<TT>(B[k]); Dc[k] = static_cast<TT>(D[k]);
        }
        // ---- pair blocks (Assembly.h:282-324) for the canonical (lo, hi) orientation:
        // K_IJ with I the neighbour entry of lo and J that of hi (K_JI = K_IJ^T).
        // Work items: the tile's nd paired slots (two blocks with the same lower
        // node and the same coupling GPs, summed by one thread from one read of the
        // GP and lower-node records per GP), then its single slots. FP32 and the
        // coupled laminate take every slot singly: the second block would not fit
        // their register budgets (six waves per SIMD; 128 VGPRs) without spilling or
        // losing a resident workgroup.
        const int nd = (sizeof(TT) == 8 && LAM != LAM_COUPLED) ? tile_nd[t] : 0;
        // A work item's slot pairs and scratch rows, loaded one item ahead: the
        // loads of the next item are issued before this item's scratch stores, so
        // waiting for them never waits for those stores to complete.
        struct SlotMeta { uint32_t pr, d1, d2; };
        auto meta = [&](int it) {
            SlotMeta m{0u, 0u, 0u};
            if (it < T - nd) {
                const bool pd = it < nd;
                const int sl = pd ? 2 * it : nd + it;
                m.pr = tpairs[tb + sl];
                m.d1 = slot_dst[tb + sl];
                if (pd) {
                    m.pr |= static_cast<uint32_t>(tpairs[tb + sl + 1]) << 16;
                    m.d2 = slot_dst[tb + sl + 1];
                }
            }
            return m;
        };
        // One work item: the slot(s) of item it, whose slot pairs and scratch rows are cur.
        auto item = [&](const int it, const SlotMeta& cur) {
            const bool paired = it < nd;
            const uint32_t pr = cur.pr & 0xffffu;
            const int ua = pr >> 8, ub1 = pr & 0xff;
            const int ub2 = paired ? static_cast<int>((cur.pr >> 16) & 0xff) : ub1;
            uint32_t mask = pm[ua] & pm[ub1];
            TT acc[9], acc2[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) { acc[k] = TT(0); acc2[k] = TT(0); }
            // Per coupling GP the block is K_IJ = T0 a1^T + T1 a2^T + x b^T + b y^T + W geo I
            // with the weight folded into T0, T1, x, y (W S, W Tu);
            // terms() forms those factors, add() accumulates them into the slot's block,
            // coupling GPs in ascending order.
            struct Terms { TT T0[3], T1[3], xv[3], yv[3], a1[3], a2[3], bv[3], wgeo; };
            // Factors of the lower node alone at coupling GP wv, which a paired item
            // forms once for its two blocks: Q = W Psi_i^T A (2 x 3), so that
            // W S = Q Psi_j, and n = N grad Psi_i scaled by W, so that W geo = n . grad Psi_j
            // (8 + 2 operations per block instead of 18 + 6).
            struct ISide { TT q00, q01, q02, q10, q11, q12, n0, n1; };
            auto iside = [&](int wv, int ri) {
                ISide s;
                const TT W = Gp(G_W, wv);
                const TT pxi = R(R_PX, ri), pyi = R(R_PY, ri);
                const TT wpx = W * pxi, wpy = W * pyi;
                // Q[0][c] = W (pxi A[0][c] + pyi A[2][c]), Q[1][c] = W (pyi A[1][c] + pxi A[2][c])
                s.q00 = wpx * Ac[0]; s.q01 = wpx * Ac[1]; s.q02 = wpy * Ac[8];
                s.q10 = wpy * Ac[3]; s.q11 = wpy * Ac[4]; s.q12 = wpx * Ac[8];
                if constexpr (LAM != LAM_ORTHO) {
                    s.q00 += wpy * Ac[6]; s.q01 += wpy * Ac[7]; s.q02 += wpx * Ac[2];
                    s.q10 += wpx * Ac[6]; s.q11 += wpx * Ac[7]; s.q12 += wpy * Ac[5];
                }
                s.n0 = pxi * Gp(G_WN, wv) + pyi * Gp(G_WN + 2, wv);
                s.n1 = pxi * Gp(G_WN + 2, wv) + pyi * Gp(G_WN + 1, wv);
                return s;
            };
            // ri, rj: the records of the slot's two nodes at coupling GP wv; is: the
            // lower node's factors at wv, or null to form the block from the records alone
            auto terms = [&](int wv, int ri, int rj, Terms& t, const ISide* is) {
                const TT W = Gp(G_W, wv);
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    t.a1[r] = Gp(G_A1 + r, wv); t.a2[r] = Gp(G_A2 + r, wv); t.bv[r] = Gp(G_B + r, wv);
                }
                const TT pxi = R(R_PX, ri), pyi = R(R_PY, ri), pxj = R(R_PX, rj), pyj = R(R_PY, rj);
                const TT bi0 = R(R_BE0, ri), bi1 = R(R_BE1, ri), bi2 = R(R_BE2, ri);
                const TT bj0 = R(R_BE0, rj), bj1 = R(R_BE1, rj), bj2 = R(R_BE2, rj);
                const TT tui = R(R_TU, ri), tuj = R(R_TU, rj);      // W Tu
                TT gi[3], gj[3];
                if constexpr (GREC) {
#pragma unroll
                    for (int r = 0; r < 3; ++r) { gi[r] = R(R_G0 + r, ri); gj[r] = R(R_G0 + r, rj); }
                } else {
                    // g = F0 C0^-1 grad Psi from the GP's frame
                    const TT ci00 = Gp(G_CI, wv), ci01 = Gp(G_CI + 1, wv), ci11 = Gp(G_CI + 2, wv);
                    const TT ci0 = ci00 * pxi + ci01 * pyi, ci1 = ci01 * pxi + ci11 * pyi;
                    const TT cj0 = ci00 * pxj + ci01 * pyj, cj1 = ci01 * pxj + ci11 * pyj;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        gi[r] = t.a1[r] * ci0 + t.a2[r] * ci1;
                        gj[r] = t.a1[r] * cj0 + t.a2[r] * cj1;
                    }
                }
                // W S = W Psi_i^T A Psi_j
                TT S00, S01, S10, S11;
                if (is) {
                    S00 = is->q00 * pxj + is->q02 * pyj; S01 = is->q01 * pyj + is->q02 * pxj;
                    S10 = is->q10 * pxj + is->q12 * pyj; S11 = is->q11 * pyj + is->q12 * pxj;
                } else {
                    // A Psi_j (3x2)
                    TT ap00 = Ac[0] * pxj, ap01 = Ac[1] * pyj, ap10 = Ac[3] * pxj, ap11 = Ac[4] * pyj;
                    TT ap20 = Ac[8] * pyj, ap21 = Ac[8] * pxj;
                    if constexpr (LAM != LAM_ORTHO) {
                        ap00 += Ac[2] * pyj; ap01 += Ac[2] * pxj; ap10 += Ac[5] * pyj; ap11 += Ac[5] * pxj;
                        ap20 += Ac[6] * pxj; ap21 += Ac[7] * pyj;
                    }
                    S00 = W * (pxi * ap00 + pyi * ap20); S01 = W * (pxi * ap01 + pyi * ap21);
                    S10 = W * (pyi * ap10 + pxi * ap20); S11 = W * (pyi * ap11 + pxi * ap21);
                }
                // z = beta_i . D beta_j + (b.v) g_i . g_j
                TT db0 = Dc[0] * bj0 + Dc[1] * bj1, db1 = Dc[3] * bj0 + Dc[4] * bj1, db2 = Dc[8] * bj2;
                if constexpr (LAM != LAM_ORTHO) {
                    db0 += Dc[2] * bj2; db1 += Dc[5] * bj2; db2 += Dc[6] * bj0 + Dc[7] * bj1;
                }
                const TT Wz = W * (bi0 * db0 + bi1 * db1 + bi2 * db2
                                   + Gp(G_BDV, wv) * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]));
                if (is) t.wgeo = is->n0 * pxj + is->n1 * pyj;
                else t.wgeo = pxi * (Gp(G_WN, wv) * pxj + Gp(G_WN + 2, wv) * pyj)
                              + pyi * (Gp(G_WN + 2, wv) * pxj + Gp(G_WN + 1, wv) * pyj);
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    t.xv[r] = -(tui * gj[r]);
                    t.yv[r] = Wz * t.bv[r] - tuj * gi[r];
                    t.T0[r] = t.a1[r] * S00 + t.a2[r] * S10;
                    t.T1[r] = t.a1[r] * S01 + t.a2[r] * S11;
                }
                if constexpr (LAM == LAM_COUPLED) {
                    const TT bbj0 = Bc[0] * bj0 + Bc[1] * bj1 + Bc[2] * bj2;
                    const TT bbj1 = Bc[3] * bj0 + Bc[4] * bj1 + Bc[5] * bj2;
                    const TT bbj2 = Bc[6] * bj0 + Bc[7] * bj1 + Bc[8] * bj2;
                    const TT bbi0 = Bc[0] * bi0 + Bc[1] * bi1 + Bc[2] * bi2;
                    const TT bbi1 = Bc[3] * bi0 + Bc[4] * bi1 + Bc[5] * bi2;
                    const TT bbi2 = Bc[6] * bi0 + Bc[7] * bi1 + Bc[8] * bi2;
                    const TT m0 = W * (pxi * bbj0 + pyi * bbj2), m1 = W * (pyi * bbj1 + pxi * bbj2);
                    const TT q0 = W * (pxj * bbi0 + pyj * bbi2), q1 = W * (pyj * bbi1 + pxj * bbi2);
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        t.xv[r] += t.a1[r] * m0 + t.a2[r] * m1;
                        t.yv[r] += t.a1[r] * q0 + t.a2[r] * q1;
                    }
                }
            };
            auto add = [&](TT (&a)[9], const Terms& t) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        TT v = a[3 * c + r] + t.T0[r] * t.a1[c] + t.T1[r] * t.a2[c] + t.xv[r] * t.bv[c]
                               + t.bv[r] * t.yv[c];
                        if (r == c) v += t.wgeo;
                        a[3 * c + r] = v;
                    }
            };
            // The record indices of the next coupling GP are read one iteration
            // ahead, so that an iteration waits for one LDS round trip (its
            // records), not two (indices, then records).
            if (paired) {
                // one coupling GP per iteration, both blocks
                int w0 = mask ? __ffs(mask) - 1 : 0;
                int li = loc[w0 * U + ua], lj = loc[w0 * U + ub1], lk = loc[w0 * U + ub2];
                while (mask) {
                    const int wc = w0, ci = li, cj = lj, ck = lk;
                    mask &= mask - 1;
                    if (mask) {
                        w0 = __ffs(mask) - 1;
                        li = loc[w0 * U + ua]; lj = loc[w0 * U + ub1]; lk = loc[w0 * U + ub2];
                    }
                    Terms t0, t1;
                    const ISide is = iside(wc, ci);
                    terms(wc, ci, cj, t0, &is);
                    terms(wc, ci, ck, t1, &is);
                    add(acc, t0);
                    add(acc2, t1);
                }
            } else {
                int w0 = mask ? __ffs(mask) - 1 : 0;
                int li = loc[w0 * U + ua], lj = loc[w0 * U + ub1];
                while (mask) {
                    const int wc = w0, ci = li, cj = lj;
                    mask &= mask - 1;
                    if (mask) {
                        w0 = __ffs(mask) - 1;
                        li = loc[w0 * U + ua]; lj = loc[w0 * U + ub1];
                    }
                    Terms t0;
                    terms(wc, ci, cj, t0, nullptr);
                    add(acc, t0);
                }
            }
            TT* dst = scr + 9 * static_cast<int64_t>(cur.d1);
#pragma unroll
            for (int k = 0; k < 9; ++k) dst[k] = acc[k];
            if (paired) {
                TT* dst2 = scr + 9 * static_cast<int64_t>(cur.d2);
#pragma unroll
                for (int k = 0; k < 9; ++k) dst2[k] = acc2[k];
            }
        };
        const int nitems = T - nd;
        if constexpr (sizeof(TT) == 8 && LAM != LAM_COUPLED) {
            // Work items go to the waves in chunks of 64 (one item per lane). The plan
            // sorts them by decreasing cost, so dealing the chunks round-robin would
            // leave the waves that drew the cheap tail idle while the others finish:
            // wave w takes chunk w, then claims the next unclaimed chunk from an LDS
            // counter, one chunk before it needs it. Each item is still summed by one
            // lane in GP order, so the blocks do not depend on which wave takes which
            // chunk. (The FP32 tangent and the coupled laminate, whose items are all
            // single slots, keep the round-robin deal: the claim loop would take the
            // FP32 kernel past its 80-register budget and the coupled one past 128.)
            const int nchunks = (nitems + 63) >> 6;
            const int lane = tid & 63;
            auto claim = [&]() {
                int k = 0;
                if (lane == 0) k = atomicAdd(chunk_next, 1);
                return __builtin_amdgcn_readfirstlane(k);   // every lane of the wave is active here
            };
            int ch = __builtin_amdgcn_readfirstlane(tid >> 6);
            SlotMeta cur = meta(ch * 64 + lane);
            int chn = ch < nchunks ? claim() : nchunks;
            while (ch < nchunks) {
                const SlotMeta nxt = meta(chn * 64 + lane);
                const int it = ch * 64 + lane;
                if (it < nitems) item(it, cur);
                ch = chn; cur = nxt;
                if (ch < nchunks) chn = claim();
            }
        } else {
            SlotMeta cur = meta(tid);
            for (int it = tid; it < nitems; it += NT) {
                const SlotMeta nxt = meta(it + NT);
                item(it, cur);
                cur = nxt;
            }
        }
    }
    // ---- residual per union node, GPs in order ---------------------------------------
    if (tid < U) {
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        for (int wv = 0; wv < ngp; ++wv) {
            const int r = loc[wv * U + tid];
            if (r == NO_REC) continue;
            f0 += RF(R_F0, r); f1 += RF(R_F1, r); f2 += RF(R_F2, r);
        }
        double* dst = fscr + 3 * static_cast<int64_t>(ub + tid);
        dst[0] = f0; dst[1] = f1; dst[2] = f2;
    }
}

template <int LAM, typename TT, bool PERSIST, bool WIDE, int NT = TILE_THREADS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu((sizeof(TT) == 4 && LAM != LAM_COUPLED && !WIDE && !PERSIST) ? 6 : 1))) void k_tile(
    const int* __restrict__ tile_gp, const int* __restrict__ tile_uoff, const int* __restrict__ tile_unodes,
    const int64_t* __restrict__ tile_toff, const uint16_t* __restrict__ tpairs, const int* __restrict__ tile_nd,
    const uint32_t* __restrict__ slot_dst,
    const int* __restrict__ off, const uint8_t* __restrict__ uloc, const uint8_t* __restrict__ urank,
    const double* __restrict__ sf, int64_t npairs, const double* __restrict__ w,
    const double* __restrict__ xu, const double* __restrict__ kbar, MatC mat, int rec_cap, int umax,
    int with_tangent, TT* __restrict__ scr, double* __restrict__ fscr, int ntiles) {
    extern __shared__ double smem[];
    // one tile per workgroup, or (PERSIST) a grid striding over the tiles
    if constexpr (PERSIST) {
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            tile_body<LAM, TT, WIDE, NT>(tile_gp, tile_uoff, tile_unodes, tile_toff, tpairs, tile_nd, slot_dst, off, uloc, urank, sf, npairs, w, xu,
                             kbar, mat, rec_cap, umax, with_tangent, scr, fscr, t, smem);
            __syncthreads();   // the next tile reuses the LDS
        }
    } else {
        tile_body<LAM, TT, WIDE, NT>(tile_gp, tile_uoff, tile_unodes, tile_toff, tpairs, tile_nd, slot_dst, off, uloc, urank, sf, npairs, w, xu,
                             kbar, mat, rec_cap, umax, with_tangent, scr, fscr, xcd_tile(blockIdx.x, gridDim.x), smem);
    }
}

// Tiles of several evaluations in one launch (serving: the requests admitted
// together). Each job is one evaluation's plan, inputs and workspace, its tiles
// numbered from tile0 on; the jobs share the laminate class, the tangent type
// and the neighbour width, and the launch's LDS is the largest job's. The job
// table is a kernel argument (passed by value, about 3 KB): its fields are read
// like k_tile's arguments, so the kernel needs no more registers than k_tile
// (127 VGPRs FP64, 79 FP32, no spills; read from a device table they took 135,
// i.e. three waves per SIMD instead of four, and FP32 spilled) and a launch
// needs no table copy.
struct TileJob {
    const int* tile_gp; const int* tile_uoff; const int* tile_unodes; const int64_t* tile_toff;
    const uint16_t* tpairs; const int* tile_nd; const uint32_t* slot_dst;
    const int* off; const uint8_t* uloc; const uint8_t* urank; const double* sf; int64_t npairs;
    const double* w; const double* xu; const double* kbar; MatC mat;
    int rec_cap, umax, with_tangent, tile0;
    void* scr; double* fscr;
};
struct TileJobTable {
    TileJob job[kTileJobsMax];
    int n;
};

template <int LAM, typename TT, bool WIDE>
__global__ __launch_bounds__(TILE_THREADS) __attribute__((amdgpu_waves_per_eu((sizeof(TT) == 4 && LAM != LAM_COUPLED && !WIDE) ? 6 : 1))) void k_tile_multi(const TileJobTable tab) {
    extern __shared__ double smem[];
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    int j = 0;
    while (j + 1 < tab.n && tab.job[j + 1].tile0 <= b) ++j;
    const TileJob& J = tab.job[j];
    tile_body<LAM, TT, WIDE, TILE_THREADS>(J.tile_gp, J.tile_uoff, J.tile_unodes, J.tile_toff, J.tpairs, J.tile_nd, J.slot_dst,
                             J.off, J.uloc, J.urank, J.sf, J.npairs, J.w, J.xu, J.kbar, J.mat, J.rec_cap, J.umax,
                             J.with_tangent, static_cast<TT*>(J.scr), J.fscr, b - J.tile0, smem);
}

// ---------------------------------------------------------------------------
// Elastic hinge per crease GP (Constraints.h:71-119, :341-436). Writes the
// GP's W*Kt and W*moment, and per node of L u R the aggregated gradient
// G(node) = sum over its occurrences of dtheta/du (3 dofs).
// ---------------------------------------------------------------------------
struct FoldDev {
    const double* lsf; const double* rsf;
    const int* loff; const int* roff;
    const int* lids; const int* rids;
    int64_t lnp, rnp;
    const double* w;       // left-domain weights
    const double* tx; const double* ty;
    const double* theta0;  // may be null
    double Kt, Ang0;
};

// Per crease GP record: [W Kt, W moment, H (12 x 12, symmetrised FD Hessian of
// theta, only with the exact tangent)]. Per (crease GP, node of L u R) record:
// G (aggregated dtheta/du, 3) and the summed shape derivatives of the node's
// left and right occurrences (dX_L, dY_L, dX_R, dY_R), which give the node's
// 12-vector map B(node, d) used by the Hessian term.
constexpr int FGP_REC = 2 + 144;
constexpr int FND_REC = 7;

// One wave (workgroup) per crease GP: both one-sided F0 by wave sums over the
// GP's neighbours, the angle and its gradient (every lane), the FD Hessian
// columns on lanes 0-11, then the per-node aggregated gradients, the left
// side's occurrences before the right side's (a node seen from both sides gets
// its left term first).
__global__ __launch_bounds__(64) void k_fold(int nfgp, const FoldDev* __restrict__ folds, const int* __restrict__ fgp_fold,
                       const int* __restrict__ fgp_local, const int* __restrict__ fgp_noff,
                       const int* __restrict__ lpo, const int* __restrict__ rpo,
                       const int* __restrict__ lpos, const int* __restrict__ rpos,
                       const double* __restrict__ xu, int exact, double* __restrict__ fgp_rec,
                       double* __restrict__ Gnd) {
    __shared__ double Hs[144];
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    if (q >= nfgp) return;                    // whole workgroup
    const FoldDev fd = folds[fgp_fold[q]];
    const int i = fgp_local[q];
    double F[12];
    // F0 of the left and right GPs (Kinematics.h:75-91)
    for (int side = 0; side < 2; ++side) {
        const double* sf = side ? fd.rsf : fd.lsf;
        const int* of = side ? fd.roff : fd.loff;
        const int* id = side ? fd.rids : fd.lids;
        const int64_t np = side ? fd.rnp : fd.lnp;
        double p1[3] = {0.0, 0.0, 0.0}, p2[3] = {0.0, 0.0, 0.0};
        for (int k = of[i] + lane; k < of[i + 1]; k += 64) {
            const int v = id[k];
            const double px = sf[np + k], py = sf[2 * np + k];
#pragma unroll
            for (int d = 0; d < 3; ++d) { p1[d] += xu[3 * v + d] * px; p2[d] += xu[3 * v + d] * py; }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            F[6 * side + d] = (d == 0 ? 1.0 : 0.0) + wave_sum(p1[d]);
            F[6 * side + 3 + d] = (d == 1 ? 1.0 : 0.0) + wave_sum(p2[d]);
        }
    }
    double grad[12];
    const double tx = fd.tx[i], ty = fd.ty[i];
    const double theta = fold_angle_grad(F, tx, ty, grad);
    const double th0 = fd.theta0 ? fd.theta0[i] : fd.Ang0;
    const double W = fd.w[i];
    double* rec = fgp_rec + static_cast<int64_t>(FGP_REC) * q;
    const double Wm = W * (fd.Kt * (theta - th0));
    if (lane == 0) { rec[0] = W * fd.Kt; rec[1] = Wm; }
    if (exact && fabs(Wm) > 0.0) {
        // Constraints.h:402-416: central differences of the analytic gradient,
        // step 1e-6 max(1, |F|_inf), then symmetrised; lane k < 12 differentiates
        // along F_k
        double fmax = 0.0;
        for (int k = 0; k < 12; ++k) fmax = fmax > fabs(F[k]) ? fmax : fabs(F[k]);
        const double hs = 1e-6 * (fmax > 1.0 ? fmax : 1.0);
        if (lane < 12) {
            const int k = lane;
            double Fp[12], Fm[12], gp[12], gm[12];
            for (int r = 0; r < 12; ++r) { Fp[r] = F[r]; Fm[r] = F[r]; }
            Fp[k] += hs; Fm[k] -= hs;
            fold_angle_grad(Fp, tx, ty, gp);
            fold_angle_grad(Fm, tx, ty, gm);
            for (int r = 0; r < 12; ++r) Hs[12 * r + k] = (gp[r] - gm[r]) / (2.0 * hs);
        }
        __syncthreads();
        for (int e = lane; e < 144; e += 64) {
            const int r = e / 12, k = e % 12;
            rec[2 + e] = (r == k) ? Hs[e] : 0.5 * (Hs[12 * r + k] + Hs[12 * k + r]);
        }
    } else {
        for (int e = lane; e < 144; e += 64) rec[2 + e] = 0.0;
    }
    const int nb0 = fgp_noff[q], nb1 = fgp_noff[q + 1];
    for (int k = nb0 + lane; k < nb1; k += 64)
        for (int e = 0; e < FND_REC; ++e) Gnd[static_cast<int64_t>(FND_REC) * k + e] = 0.0;
    for (int side = 0; side < 2; ++side) {
        __syncthreads();   // the previous writes of the GP's node records are visible
        const double* sf = side ? fd.rsf : fd.lsf;
        const int* of = side ? fd.roff : fd.loff;
        const int64_t np = side ? fd.rnp : fd.lnp;
        const int* pos = side ? (rpos + rpo[q]) : (lpos + lpo[q]);
        const int blk = 6 * side;
        // one side's neighbours are distinct nodes: no two lanes share a record
        for (int k = of[i] + lane; k < of[i + 1]; k += 64) {
            const double dX = sf[np + k], dY = sf[2 * np + k];
            double* G = Gnd + static_cast<int64_t>(FND_REC) * (nb0 + pos[k - of[i]]);
#pragma unroll
            for (int d = 0; d < 3; ++d) G[d] += grad[blk + d] * dX + grad[blk + d + 3] * dY;
            G[3 + 2 * side] += dX;
            G[4 + 2 * side] += dY;
        }
    }
}

// ---------------------------------------------------------------------------
// Segmented reductions (one thread per K block / per node).
// ---------------------------------------------------------------------------
constexpr int RK_THREADS = 256;   // K blocks per workgroup of k_reduce_K
constexpr int RK_UNITS = 8;       // 16-byte units per thread per staged chunk
typedef unsigned int rk_unit __attribute__((ext_vector_type(4)));   // one 16-byte unit
// rows per chunk: 16 * RK_UNITS * RK_THREADS bytes of LDS hold the chunk's 9 n
// elements plus the up to U - 1 elements before it in its first 16-byte unit
template <typename TT> constexpr int rk_chunk() {
    return (RK_UNITS * RK_THREADS * static_cast<int>(16 / sizeof(TT)) - static_cast<int>(16 / sizeof(TT) - 1)) / 9;
}

// The block's fold terms added to its face sum vt, and the 3x3 block written
// into the CSC of K.
template <typename TT>
__device__ __forceinline__ void finish_K_block(int64_t blk, int64_t bs, const TT (&vt)[9],
                                               const int64_t* __restrict__ kf_off, const int* __restrict__ kf_idx,
                                               const int* __restrict__ nd_gp, const double* __restrict__ fgp_rec,
                                               const double* __restrict__ Gnd, int folds, int exact,
                                               const int* __restrict__ stride, double* __restrict__ K) {
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = static_cast<double>(vt[k]);
    if (folds) {
        double f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k] = 0.0;
        for (int64_t a = kf_off[blk]; a < kf_off[blk + 1]; ++a) {
            const int lo = kf_idx[2 * a], hi = kf_idx[2 * a + 1];
            const double* rec = fgp_rec + static_cast<int64_t>(FGP_REC) * nd_gp[lo];
            const double wk = rec[0];
            const double* Gl = Gnd + static_cast<int64_t>(FND_REC) * lo;
            const double* Gh = Gnd + static_cast<int64_t>(FND_REC) * hi;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) f[3 * c + r] += Gl[r] * Gh[c] * wk;
            if (exact) {
                // W moment * B(lo, r)^T H B(hi, c); B(node, d) has dX_L, dY_L, dX_R, dY_R
                // at positions d, d + 3, 6 + d, 9 + d of the 12-vector
                const double wm = rec[1];
                const double* H = rec + 2;
                const double bl[4] = {Gl[3], Gl[4], Gl[5], Gl[6]};
                const double bh[4] = {Gh[3], Gh[4], Gh[5], Gh[6]};
                // one column c of B(lo)^T H B(hi) per iteration (not unrolled, so that
                // the 144 values of H are not all held at once)
#pragma unroll 1
                for (int c = 0; c < 3; ++c) {
                    // hv[p] = sum_v H(p, qc[v]) bh[v] with qc = {c, c + 3, 6 + c, 9 + c}
                    double hv[12];
#pragma unroll
                    for (int p = 0; p < 12; ++p)
                        hv[p] = H[12 * p + c] * bh[0] + H[12 * p + c + 3] * bh[1] + H[12 * p + 6 + c] * bh[2]
                                + H[12 * p + 9 + c] * bh[3];
                    double col[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        col[r] = wm * (bl[0] * hv[r] + bl[1] * hv[r + 3] + bl[2] * hv[6 + r] + bl[3] * hv[9 + r]);
#pragma unroll
                    for (int k = 0; k < 9; ++k) f[k] += (k / 3 == c) ? col[k % 3] : 0.0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] += f[k];
    }
    const int st = stride[blk];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) K[bs + static_cast<int64_t>(c) * st + r] = v[3 * c + r];
}


// A workgroup owns RK_THREADS consecutive K blocks, whose scratch rows are one
// contiguous range. The range is staged through LDS in chunks of 16-byte units
// (consecutive lanes, consecutive units; the next chunk's loads are issued
// before the current chunk is summed); each thread adds the rows of its own
// block from LDS in ascending order, the same order and rounding as a direct
// walk over the rows. The scratch buffer carries 16 bytes of padding, so the
// last unit of the last chunk stays inside it.
template <typename TT>
__global__ __launch_bounds__(RK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_reduce_K(int64_t nblocks, const int64_t* __restrict__ kc_off,
                           const TT* __restrict__ scr, const int64_t* __restrict__ kf_off,
                           const int* __restrict__ kf_idx, const int* __restrict__ nd_gp,
                           const double* __restrict__ fgp_rec, const double* __restrict__ Gnd, int folds,
                           int exact, const int64_t* __restrict__ base, const int* __restrict__ stride,
                           double* __restrict__ K) {
    constexpr int U = 16 / sizeof(TT);              // elements per unit
    constexpr int CH = rk_chunk<TT>();
    __shared__ rk_unit lds_units[RK_UNITS * RK_THREADS];
    const TT* rows = reinterpret_cast<const TT*>(lds_units);
    const int tid = threadIdx.x;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * RK_THREADS;
    const int64_t b1 = b0 + RK_THREADS < nblocks ? b0 + RK_THREADS : nblocks;
    const int64_t blk = b0 + tid;
    const bool own = blk < nblocks;
    const int64_t r0 = kc_off[b0], r1 = kc_off[b1];
    const int64_t a0 = own ? kc_off[blk] : 0, a1 = own ? kc_off[blk + 1] : 0;
    // chunk at row c: units [9 c / U, ceil(9 min(r1, c + CH) / U)) of the scratch
    auto units = [&](int64_t c, int64_t& u0) {
        u0 = 9 * c / U;
        const int64_t e1 = 9 * (r1 < c + CH ? r1 : c + CH);
        return static_cast<int>((e1 + U - 1) / U - u0);
    };
    const rk_unit* src = reinterpret_cast<const rk_unit*>(scr);
    // the chunk's units go to registers first (every lane loads RK_UNITS units,
    // indices clamped to the chunk's last unit), then to LDS
    rk_unit buf[RK_UNITS];
    int64_t fu0 = 0;
    int fnu = 0;
    if (r0 < r1) {
        fnu = units(r0, fu0);
#pragma unroll
        for (int j = 0; j < RK_UNITS; ++j) {
            const int u = tid + j * RK_THREADS;
            buf[j] = src[fu0 + (u < fnu ? u : fnu - 1)];
        }
    }
    TT vt[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) vt[k] = TT(0);
    for (int64_t c = r0; c < r1; c += CH) {
        const int64_t u0 = fu0;
        const int nu = fnu;
#pragma unroll
        for (int j = 0; j < RK_UNITS; ++j) {
            const int u = tid + j * RK_THREADS;
            if (u < nu) lds_units[u] = buf[j];
        }
        __syncthreads();
        if (c + CH < r1) {
            fnu = units(c + CH, fu0);
#pragma unroll
            for (int j = 0; j < RK_UNITS; ++j) {
                const int u = tid + j * RK_THREADS;
                buf[j] = src[fu0 + (u < fnu ? u : fnu - 1)];
            }
        }
        const int shift = static_cast<int>(9 * c - u0 * U);
        const int64_t lo = a0 > c ? a0 : c, hi = a1 < c + CH ? a1 : c + CH;
        for (int64_t a = lo; a < hi; ++a) {
            const TT* s = rows + shift + 9 * (a - c);
#pragma unroll
            for (int k = 0; k < 9; ++k) vt[k] += s[k];
        }
        __syncthreads();
    }
    if (!own) return;
    const int64_t bs = base[blk];
    if (bs < 0) return;
    finish_K_block(blk, bs, vt, kf_off, kf_idx, nd_gp, fgp_rec, Gnd, folds, exact, stride, K);
}

// Direct walk: one lane per block adds its contiguous scratch rows in
// ascending order (the staged kernel's order and rounding). Used when blocks
// have many rows (large supports), where the staged kernel would leave most
// lanes of a chunk idle while a few walk long row ranges.
template <typename TT>
__global__ __launch_bounds__(RK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_reduce_K_rows(int64_t nblocks, const int64_t* __restrict__ kc_off,
                           const TT* __restrict__ scr, const int64_t* __restrict__ kf_off,
                           const int* __restrict__ kf_idx, const int* __restrict__ nd_gp,
                           const double* __restrict__ fgp_rec, const double* __restrict__ Gnd, int folds,
                           int exact, const int64_t* __restrict__ base, const int* __restrict__ stride,
                           double* __restrict__ K) {
    const int64_t blk = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (blk >= nblocks) return;
    const int64_t bs = base[blk];
    if (bs < 0) return;
    TT vt[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) vt[k] = TT(0);
    const int64_t a0 = kc_off[blk], a1 = kc_off[blk + 1];
#pragma unroll 4
    for (int64_t a = a0; a < a1; ++a) {
        const TT* s = scr + 9 * a;
#pragma unroll
        for (int k = 0; k < 9; ++k) vt[k] += s[k];
    }
    finish_K_block(blk, bs, vt, kf_off, kf_idx, nd_gp, fgp_rec, Gnd, folds, exact, stride, K);
}

__global__ void k_reduce_F(int nn, const int* __restrict__ fc_off, const int* __restrict__ fc_idx,
                           const double* __restrict__ fscr, const int* __restrict__ ff_off,
                           const int* __restrict__ ff_idx, const int* __restrict__ nd_gp,
                           const double* __restrict__ fgp_rec, const double* __restrict__ Gnd, int folds,
                           double* __restrict__ F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int a = fc_off[i]; a < fc_off[i + 1]; ++a) {
        const double* s = fscr + 3 * static_cast<int64_t>(fc_idx[a]);
        f0 += s[0]; f1 += s[1]; f2 += s[2];
    }
    if (folds) {
        for (int a = ff_off[i]; a < ff_off[i + 1]; ++a) {
            const int k = ff_idx[a];
            const double wm = fgp_rec[static_cast<int64_t>(FGP_REC) * nd_gp[k] + 1];
            const double* G = Gnd + static_cast<int64_t>(FND_REC) * k;
            f0 += G[0] * wm; f1 += G[1] * wm; f2 += G[2] * wm;
        }
    }
    F[3 * i] = f0; F[3 * i + 1] = f1; F[3 * i + 2] = f2;
}

// k_tile instantiation: body couplings, persistent grid, GPs with more than 64
// neighbours (chunked lanes)
template <typename TT, bool PERSIST, bool WIDE>
static auto tile_kernel(int lam) {
    return lam == LAM_COUPLED ? k_tile<LAM_COUPLED, TT, PERSIST, WIDE>
         : lam == LAM_SYM     ? k_tile<LAM_SYM, TT, PERSIST, WIDE>
                              : k_tile<LAM_ORTHO, TT, PERSIST, WIDE>;
}

template <typename TT>
static auto tile_kernel(int lam, bool persist, bool wide) {
    return persist ? (wide ? tile_kernel<TT, true, true>(lam) : tile_kernel<TT, true, false>(lam))
                   : (wide ? tile_kernel<TT, false, true>(lam) : tile_kernel<TT, false, false>(lam));
}

// The FP32 tangent over wide tiles whose LDS leaves room for one workgroup per
// CU: 1024 threads (four waves per SIMD) instead of 512 (two). Its 98 VGPRs fit
// four waves; the FP64 kernel at 1024 threads would spill past 128 VGPRs. At the
// reference defaults (cubic basis, rho = 6h) the FP32 tile takes 15.9 ms instead
// of 17.9 ms with bit-identical results.
static auto tile_kernel_fp32_wide_big(int lam) {
    constexpr int NT = 2 * TILE_THREADS;
    return lam == LAM_COUPLED ? k_tile<LAM_COUPLED, float, false, true, NT>
         : lam == LAM_SYM     ? k_tile<LAM_SYM, float, false, true, NT>
                              : k_tile<LAM_ORTHO, float, false, true, NT>;
}

// The FP64 tangent (and the residual alone) in the same case: 768 threads, three
// waves per SIMD, within the 170 VGPRs that three waves leave each (22.4 ms
// instead of 23.5 ms at the reference defaults, bit-identical).
static auto tile_kernel_fp64_wide_big(int lam) {
    constexpr int NT = 3 * TILE_THREADS / 2;
    return lam == LAM_COUPLED ? k_tile<LAM_COUPLED, double, false, true, NT>
         : lam == LAM_SYM     ? k_tile<LAM_SYM, double, false, true, NT>
                              : k_tile<LAM_ORTHO, double, false, true, NT>;
}

int tile_umax(int max_union) { return std::min(UMAX, (std::max(max_union, 1) + 7) / 8 * 8); }

size_t tile_lds_bytes(int rec_cap, int max_union, bool fp32, bool wide) {
    const size_t tt = fp32 ? sizeof(float) : sizeof(double);
    const size_t um = static_cast<size_t>(tile_umax(max_union));
    return sizeof(double) * (static_cast<size_t>(rec_cap) * 3 + um * 3)
         + tt * (static_cast<size_t>(rec_cap) * (fp32 ? (wide ? rec_stride<float, false>() : rec_stride<float, true>())
                                                      : (wide ? rec_stride<double, false>() : rec_stride<double, true>())) +
                 TILE_GPS * (fp32 ? gp_stride<float>() : gp_stride<double>())) + sizeof(uint32_t) * um
         + sizeof(uint16_t) * TILE_GPS * um + sizeof(int);
}

size_t assembly_workspace_bytes(const Context& c, int flags) {
    const Plan& P = c.plan;
    const bool tangent = (flags & CVM_TANGENT) != 0;
    const bool fp32 = (flags & CVM_FP32_TANGENT) != 0;
    size_t b = sizeof(double) * std::max<int64_t>(3 * P.nfscratch, 1);
    if (tangent) b += (fp32 ? sizeof(float) : sizeof(double)) * 9 * P.nscratch + 16;
    b += sizeof(double) * (static_cast<size_t>(FGP_REC) * P.nfgp + static_cast<size_t>(FND_REC) *
                           std::max<int64_t>(P.nfnodes, 1));
    return b;
}

// Element counts of the four workspace buffers: fscr, scr (0 without the
// tangent), fold_gp, fold_nd.
void workspace_counts(const Plan& P, int flags, size_t (&n)[4]) {
    n[0] = static_cast<size_t>(std::max<int64_t>(3 * P.nfscratch, 1));
    n[1] = 0;
    if (flags & CVM_TANGENT) {
        const bool fp32 = (flags & CVM_FP32_TANGENT) != 0;
        // + 16 bytes: k_reduce_K reads the scratch in whole 16-byte units
        n[1] = static_cast<size_t>((fp32 ? (9 * P.nscratch + 1) / 2 : 9 * P.nscratch) + 2);
    }
    n[2] = static_cast<size_t>(std::max<int64_t>(static_cast<int64_t>(FGP_REC) * P.nfgp, 1));
    n[3] = static_cast<size_t>(FND_REC) * static_cast<size_t>(std::max<int64_t>(P.nfnodes, 1));
}

void reserve_workspace(const Context& c, Workspace& ws, int flags) {
    size_t n[4];
    workspace_counts(c.plan, flags, n);
    ws.fscr.ensure(n[0]);
    if (n[1]) ws.scr.ensure(n[1]);
    ws.fold_gp.ensure(n[2]);
    ws.fold_nd.ensure(n[3]);
}

bool workspace_would_grow(const Context& c, const Workspace& ws, int flags) {
    size_t n[4];
    workspace_counts(c.plan, flags, n);
    auto grows = [](const auto& b, size_t k) { return k > b.n || (k && !b.p); };
    return grows(ws.fscr, n[0]) || (n[1] && grows(ws.scr, n[1])) || grows(ws.fold_gp, n[2]) ||
           grows(ws.fold_nd, n[3]);
}

// A kernel launch, or (sig->dry) only its contribution to the launch signature.
template <typename K, typename... A>
static void launch_k(LaunchSig* sig, K kern, dim3 grid, dim3 block, size_t lds, hipStream_t st, A... args) {
    if (sig) {
        const void* kp = reinterpret_cast<const void*>(kern);
        sig->mix(&kp, sizeof(kp));
        const unsigned dims[6] = {grid.x, grid.y, grid.z, block.x, block.y, block.z};
        sig->mix(dims, sizeof(dims));
        sig->mix(&lds, sizeof(lds));
        (sig->mix(&args, sizeof(args)), ...);
        if (sig->dry) return;
    }
    hipLaunchKernelGGL(kern, grid, block, lds, st, args...);
}

const double* natural_kappa(Context& c, bool allow_alloc) {
    if (!c.nat.active) return nullptr;
    Domain* fd = nullptr;
    for (Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) fd = d;
    if (!fd) throw Error(CVM_E_STATE, "natural curvature needs a face domain");
    if (c.kbar_ng == fd->ng && c.d_kbar.p) return c.d_kbar.p;
    if (!allow_alloc) throw Error(CVM_E_STATE, "natural curvature field not formed (cvm_init_assembly after the face GPs change)");
    std::vector<double> k(std::max(fd->ng, 1), 0.0);
    for (int g = 0; g < fd->ng; ++g) k[g] = c.nat.kappa00(fd->hx[g]);
    c.d_kbar.upload(k, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    c.kbar_ng = fd->ng;
    return c.d_kbar.p;
}

static int laminate_class(const double* A, const double* B, const double* D) {
    auto zero = [](const double* a, std::initializer_list<int> ks) {
        for (int k : ks) if (a[k] != 0.0) return false;
        return true;
    };
    return !zero(B, {0, 1, 2, 3, 4, 5, 6, 7, 8}) ? LAM_COUPLED
         : zero(A, {2, 5, 6, 7}) && zero(D, {2, 5, 6, 7}) ? LAM_ORTHO : LAM_SYM;
}

int tile_group_key(const Context& c, int flags) {
    const Plan& P = c.plan;
    if (!P.ready || (c.tile_grid > 0 && c.tile_grid < P.ntiles)) return -1;
    return ((flags & CVM_FP32_TANGENT) ? 1 : 0) | (laminate_class(c.A, c.B, c.D) << 1) | ((P.max_nbr > 64 ? 1 : 0) << 3);
}

template <typename TT, bool WIDE>
static auto tile_multi_kernel(int lam) {
    return lam == LAM_COUPLED ? k_tile_multi<LAM_COUPLED, TT, WIDE>
         : lam == LAM_SYM     ? k_tile_multi<LAM_SYM, TT, WIDE>
                              : k_tile_multi<LAM_ORTHO, TT, WIDE>;
}

void launch_tile_group(const TileGroupItem* items, int n, hipStream_t st) {
    if (n <= 0) return;
    if (n > kTileJobsMax) throw Error(CVM_E_INVALID, "more evaluations than one tile launch takes");
    TileJobTable tab;
    std::memset(&tab, 0, sizeof(tab));
    tab.n = n;
    TileJob* H = tab.job;
    const int key = tile_group_key(*items[0].c, items[0].flags);
    if (key < 0) throw Error(CVM_E_STATE, "evaluation cannot share a tile launch");
    size_t lds = 0;
    int tiles = 0;
    for (int i = 0; i < n; ++i) {
        Context& c = *items[i].c;
        if (tile_group_key(c, items[i].flags) != key) throw Error(CVM_E_STATE, "grouped evaluations differ in kind");
        const Plan& P = c.plan;
        Domain* fdp = nullptr;
        for (Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) fdp = d;
        const Domain& fd = *fdp;
        TileJob& J = H[i];
        J.tile_gp = P.d_tile_gp.p; J.tile_uoff = P.d_tile_uoff.p; J.tile_unodes = P.d_tile_unodes.p;
        J.tile_toff = P.d_tile_toff.p; J.tpairs = P.d_tile_tpairs.p; J.tile_nd = P.d_tile_nd.p;
        J.slot_dst = P.d_slot_dst.p; J.off = fd.off.p; J.uloc = P.d_pair_uloc.p; J.urank = P.d_pair_urank.p;
        J.sf = fd.sf.p; J.npairs = static_cast<int64_t>(fd.npairs); J.w = fd.d_w.p; J.xu = items[i].x;
        J.kbar = natural_kappa(c, false);
        std::memcpy(J.mat.A, c.A, sizeof(J.mat.A));
        std::memcpy(J.mat.B, c.B, sizeof(J.mat.B));
        std::memcpy(J.mat.D, c.D, sizeof(J.mat.D));
        J.rec_cap = P.rec_cap; J.umax = tile_umax(P.max_union);
        J.with_tangent = (items[i].flags & CVM_TANGENT) ? 1 : 0;
        J.tile0 = tiles;
        J.scr = items[i].scr; J.fscr = items[i].fscr;
        tiles += P.ntiles;
        lds = std::max(lds, tile_lds_bytes(P.rec_cap, P.max_union, (key & 1) != 0, (key >> 3) != 0));
    }
    const int lam = (key >> 1) & 3;
    const bool fp32 = (key & 1) != 0, wide = (key >> 3) != 0;
    auto kern = fp32 ? (wide ? tile_multi_kernel<float, true>(lam) : tile_multi_kernel<float, false>(lam))
                     : (wide ? tile_multi_kernel<double, true>(lam) : tile_multi_kernel<double, false>(lam));
    hipLaunchKernelGGL(kern, dim3(tiles), dim3(TILE_THREADS), lds, st, tab);
    CVM_HIP(hipGetLastError());
}

void upload_fold_descriptors(Context& c) {
    Plan& P = c.plan;
    std::vector<FoldDev> fdv(std::max<size_t>(c.folds.size(), 1));
    for (size_t f = 0; f < c.folds.size(); ++f) {
        const Fold& fo = c.folds[f];
        Domain& L = *c.doms[fo.left];
        Domain& R = *c.doms[fo.right];
        FoldDev& v = fdv[f];
        v.lsf = L.sf.p; v.rsf = R.sf.p; v.loff = L.off.p; v.roff = R.off.p; v.lids = L.ids.p; v.rids = R.ids.p;
        v.lnp = L.npairs; v.rnp = R.npairs; v.w = L.d_w.p; v.tx = L.d_tx.p; v.ty = L.d_ty.p;
        v.theta0 = fo.theta0.empty() ? nullptr : fo.d_theta0.p;
        v.Kt = fo.Kt; v.Ang0 = fo.Ang0;
    }
    P.d_folddev.upload(reinterpret_cast<const unsigned char*>(fdv.data()), fdv.size() * sizeof(FoldDev), c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
}

// One residual (+ tangent) evaluation on workspace ws. Reads only the plan and
// the shape functions, so evaluations on different workspaces may run
// concurrently on different streams.
void launch_assembly(Context& c, Workspace& ws, int flags, const double* xdev, double* Fdev, double* Kdev,
                     LaunchSig* sig, const WsPtrs* ext, bool tile_done) {
    Plan& P = c.plan;
    const bool dry = sig && sig->dry;
    if (!P.ready) throw Error(CVM_E_STATE, "assembly plan not initialised (cvm_init_assembly)");
    Domain* fdp = nullptr;
    for (Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) fdp = d;
    Domain& fd = *fdp;
    const hipStream_t st = ws.stream;
    auto tbeg = [&](const char* name) { return (ws.timed && !dry) ? timing_begin(c, name) : -1; };
    auto tend = [&](int slot) { if (ws.timed && !dry) timing_end(c, slot); };
    const bool tangent = (flags & CVM_TANGENT) != 0;
    const bool folds = (flags & CVM_ELASTIC_FOLDS) != 0 && P.nfgp > 0;
    const bool exact = folds && tangent && (flags & CVM_EXACT_HINGE) != 0;
    const bool fp32 = (flags & CVM_FP32_TANGENT) != 0;
    MatC m;
    std::memcpy(m.A, c.A, sizeof(m.A));
    std::memcpy(m.B, c.B, sizeof(m.B));
    std::memcpy(m.D, c.D, sizeof(m.D));
    if (ext) {
        // caller-provided workspace: nothing is reserved
    } else if (dry) {
        if (workspace_would_grow(c, ws, flags)) sig->grow = true;
    } else {
        reserve_workspace(c, ws, flags);
    }
    const WsPtrs wp = ext ? *ext : WsPtrs{ws.scr.p, ws.fscr.p, ws.fold_gp.p, ws.fold_nd.p};
    const bool wide = P.max_nbr > 64;
    const size_t lds = tile_lds_bytes(P.rec_cap, P.max_union, fp32, wide);
    // the field is formed when the plan is built or the band is set, so that
    // evaluations (the allocation-free one and graph captures among them) only read it
    const double* kbar = natural_kappa(c, false);
    const int umax = tile_umax(P.max_union);
    int slot = -1;
    const int lam = laminate_class(m.A, m.B, m.D);
    const bool persist = c.tile_grid > 0 && c.tile_grid < P.ntiles;
    const int grid = persist ? c.tile_grid : P.ntiles;
    // one workgroup per CU by LDS: the wide tile runs 1024 threads (FP32) or 768 (FP64)
    const bool big = wide && !persist && 2 * lds > c.lds_limit;
    const int nt = !big ? TILE_THREADS : fp32 ? 2 * TILE_THREADS : 3 * TILE_THREADS / 2;
    auto tile = [&](auto kern, auto* scr) {
        launch_k(sig, kern, dim3(grid), dim3(nt), lds, st, P.d_tile_gp.p, P.d_tile_uoff.p,
                           P.d_tile_unodes.p, P.d_tile_toff.p, P.d_tile_tpairs.p, P.d_tile_nd.p, P.d_slot_dst.p, fd.off.p,
                           P.d_pair_uloc.p, P.d_pair_urank.p, fd.sf.p, static_cast<int64_t>(fd.npairs), fd.d_w.p, xdev,
                           kbar, m, P.rec_cap, umax, tangent ? 1 : 0, scr, wp.fscr, P.ntiles);
    };
    float* scr32 = reinterpret_cast<float*>(wp.scr);
    if (!tile_done) {    // (tile_done: the tiles ran in a grouped launch, launch_tile_group)
        slot = tbeg(!tangent ? "tile_residual" : (fp32 ? "tile_residual_tangent_fp32" : "tile_residual_tangent"));
        if (big && fp32) tile(tile_kernel_fp32_wide_big(lam), scr32);
        else if (big) tile(tile_kernel_fp64_wide_big(lam), wp.scr);
        else if (fp32) tile(tile_kernel<float>(lam, persist, wide), scr32);
        else tile(tile_kernel<double>(lam, persist, wide), wp.scr);
        if (!dry) CVM_HIP(hipGetLastError());
        tend(slot);
    }
    if (folds) {
        slot = tbeg("fold_hinge");
        if (P.nfgp > 0) launch_k(sig, k_fold, dim3(P.nfgp), dim3(64), 0, st, P.nfgp,
                           reinterpret_cast<const FoldDev*>(P.d_folddev.p), P.d_fgp_fold.p, P.d_fgp_local.p,
                           P.d_fgp_noff.p, P.d_fgp_lpo.p, P.d_fgp_rpo.p, P.d_fgp_lpos.p, P.d_fgp_rpos.p, xdev,
                           exact ? 1 : 0, wp.fold_gp, wp.fold_nd);
        if (!dry) CVM_HIP(hipGetLastError());
        tend(slot);
    }
    slot = tbeg("reduce_F");
    launch_k(sig, k_reduce_F, dim3((c.nn + 255) / 256), dim3(256), 0, st, c.nn, P.d_fc_off.p, P.d_fc_idx.p,
                       wp.fscr, P.d_ff_off.p, P.d_ff_idx.p, P.d_fnd_gp.p, wp.fold_gp, wp.fold_nd,
                       folds ? 1 : 0, Fdev);
    if (!dry) CVM_HIP(hipGetLastError());
    tend(slot);
    if (tangent) {
        const bool uni = (flags & CVM_ELASTIC_FOLDS) != 0;
        slot = tbeg("reduce_K");
        auto red = [&](auto kern, auto* scr) {
            launch_k(sig, kern, dim3(static_cast<unsigned>((P.nblocks + RK_THREADS - 1) / RK_THREADS)),
                               dim3(RK_THREADS), 0, st,
                               P.nblocks, P.d_kc_off.p, scr, P.d_kf_off.p, P.d_kf_idx.p, P.d_fnd_gp.p,
                               wp.fold_gp, wp.fold_nd, folds ? 1 : 0, exact ? 1 : 0,
                               uni ? P.d_blk_base_u.p : P.d_blk_base_f.p,
                               uni ? P.d_blk_stride_u.p : P.d_blk_stride_f.p, Kdev);
        };
        // staged through LDS up to 24 scratch rows per block on average (supports
        // of ~4 h), the direct walk beyond
        const bool staged = P.nscratch <= 24 * P.nblocks;
        if (fp32) red(staged ? k_reduce_K<float> : k_reduce_K_rows<float>, scr32);
        else red(staged ? k_reduce_K<double> : k_reduce_K_rows<double>, wp.scr);
        if (!dry) CVM_HIP(hipGetLastError());
        tend(slot);
    }
}

// ---------------------------------------------------------------------------
// Allocation-free evaluation. Its workspace is this file-scope device array:
// one copy per GPU, loaded with the code object and held until the process
// ends. In host code `scratch` names a host-side stand-in, not the device
// memory, so the device address of each GPU's copy is taken once with
// hipGetSymbolAddress (cvm_create calls noalloc_scratch_init with the device
// current) and only that address is passed to kernels. Every call on one GPU
// shares the array: calls that overlap on one GPU race. Nothing in it needs
// clearing between calls: each kernel writes every scratch element that a
// later kernel of the same evaluation reads.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = static_cast<size_t>(256) << 20;
static_assert(kNoAllocScratchBytes == static_cast<size_t>(CVM_NOALLOC_SCRATCH_BYTES), "scratch size");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static hipError_t g_scratch_err[kMaxDevices] = {};

// Called with `device` current; a failure is kept and reported by the entry point.
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_scratch_ptr[device] = p;
    else g_scratch_err[device] = (e == hipSuccess) ? hipErrorInvalidSymbol : e;
}

// Byte offsets of the four workspace regions in the scratch (256-byte aligned)
// and the total.
static size_t noalloc_layout(const Plan& P, int flags, size_t (&off)[4]) {
    size_t n[4];
    workspace_counts(P, flags, n);
    size_t at = 0;
    for (int k = 0; k < 4; ++k) {
        off[k] = at;
        at += (n[k] * sizeof(double) + 255) / 256 * 256;
    }
    return at;
}

void launch_assembly_noalloc(Context& c, int flags, const double* xdev, double* Fdev, double* Kdev) {
    void* base = nullptr;
    hipError_t err = hipSuccess;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxDevices) { base = g_scratch_ptr[c.device]; err = g_scratch_err[c.device]; }
    }
    if (!base)
        throw Error(CVM_E_HIP, std::string("static scratch of this device unavailable (hipGetSymbolAddress: ") +
                               hipGetErrorString(err) + ")");
    size_t off[4];
    const size_t need = noalloc_layout(c.plan, flags, off);
    if (need > kNoAllocScratchBytes)
        throw Error(CVM_E_BUDGET, "cvm_assemble_noalloc: this sheet needs " + std::to_string(need) +
                                  " bytes of workspace; the static scratch holds " +
                                  std::to_string(kNoAllocScratchBytes) + " (use cvm_assemble)");
    unsigned char* b = static_cast<unsigned char*>(base);
    WsPtrs wp;
    wp.fscr = reinterpret_cast<double*>(b + off[0]);
    wp.scr = reinterpret_cast<double*>(b + off[1]);
    wp.fold_gp = reinterpret_cast<double*>(b + off[2]);
    wp.fold_nd = reinterpret_cast<double*>(b + off[3]);
    const bool timed = c.ws.timed;
    c.ws.timed = false;          // timing events are created lazily: none here
    try {
        launch_assembly(c, c.ws, flags, xdev, Fdev, Kdev, nullptr, &wp);
    } catch (...) {
        c.ws.timed = timed;
        throw;
    }
    c.ws.timed = timed;
}

}  // namespace cvm
