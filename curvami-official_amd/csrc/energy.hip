// Strain energies, hinge energy and crease fold angles at a displacement state:
//   Meshless::computeStrainEnergies    plate/Meshless.h:205-219
//   Meshless::computeTotalHingeEnergy  plate/Meshless.h:198-203
//   Constraints::computeHingeEnergy    plate/Constraints.h:310-339
//   the per-GP angle of Progressive::activeFoldAngles, plate/ProgressiveCreasing.h:157-179
//
// k_face_energy: one wave per face GP, lanes over its neighbours (the tile
// kernel's kinematics without the per-neighbour records), writing the GP's
// 1/2 W e0.N and 1/2 W e1.M. k_crease_angle: one thread per crease GP, its
// left / right deformation gradients and the dihedral angle, plus the GP's
// hinge energy 1/2 W Kt (theta - theta0)^2 when active. k_sum3 sums the three
// per-GP arrays in a fixed order (strided per thread, then a fixed LDS tree),
// so repeated calls give identical bits. k_export_fields evaluates the
// exporter's per-GP columns the same way (wave per GP).
#include <algorithm>

#include "cvm_internal.h"
#include "device_common.h"
#include "fold_angle.h"

namespace cvm {

constexpr int EN_THREADS = 256;   // 4 waves: 4 face GPs per block
constexpr int SUM_THREADS = 1024;

template <bool HAS_B>
__global__ __launch_bounds__(EN_THREADS) void k_face_energy(int ng, const int* __restrict__ off,
                                                             const int* __restrict__ ids,
                                                             const double* __restrict__ sf, int64_t npairs,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ xu,
                                                             const double* __restrict__ kbar, MatC mat,
                                                             double* __restrict__ em, double* __restrict__ eb) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (EN_THREADS / 64) + (threadIdx.x >> 6);
    if (g >= ng) return;   // uniform per wave
    double p[15];
    for (int k = 0; k < 15; ++k) p[k] = 0.0;
    for (int k = off[g] + lane; k < off[g + 1]; k += 64) {
        const int v = ids[k];
        const double u0 = xu[3 * v], u1 = xu[3 * v + 1], u2 = xu[3 * v + 2];
        const double d[5] = {sf[npairs + k], sf[2 * npairs + k], sf[3 * npairs + k], sf[4 * npairs + k],
                             sf[5 * npairs + k]};
#pragma unroll
        for (int q = 0; q < 5; ++q) { p[3 * q] += u0 * d[q]; p[3 * q + 1] += u1 * d[q]; p[3 * q + 2] += u2 * d[q]; }
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) p[k] = wave_sum(p[k]);
    if (lane != 0) return;
    // r0,X = e_X + u,X; r0,Y = e_Y + u,Y; r0,XX / YY / XY (Kinematics.h:66-91)
    const double a1[3] = {1.0 + p[0], p[1], p[2]}, a2[3] = {p[3], 1.0 + p[4], p[5]};
    const double* rxx = p + 6; const double* ryy = p + 9; const double* rxy = p + 12;
    const double C00 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
    const double C11 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
    const double C01 = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    double n[3];
    cross3(a1, a2, n);
    const double J = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double b[3] = {n[0] / J, n[1] / J, n[2] / J};
    // Voigt strains: e0 = (E0_11, E0_22, 2 E0_12), E0 = (C0 - I) / 2; e1 from
    // E1 = -r0,ab . b (Kinematics.h:99-107, Constitutive.h:57-70)
    const double e0[3] = {0.5 * (C00 - 1.0), 0.5 * (C11 - 1.0), 2.0 * (0.5 * C01)};
    double e1[3] = {-(rxx[0] * b[0] + rxx[1] * b[1] + rxx[2] * b[2]),
                    -(ryy[0] * b[0] + ryy[1] * b[1] + ryy[2] * b[2]),
                    2.0 * -(rxy[0] * b[0] + rxy[1] * b[1] + rxy[2] * b[2])};
    if (kbar) e1[0] += kbar[g];   // E1_el = E1 + kappa_bar (Meshless.h:212)
    double N[3], M[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        N[r] = mat.A[3 * r] * e0[0] + mat.A[3 * r + 1] * e0[1] + mat.A[3 * r + 2] * e0[2];
        M[r] = mat.D[3 * r] * e1[0] + mat.D[3 * r + 1] * e1[1] + mat.D[3 * r + 2] * e1[2];
        if (HAS_B) {
            N[r] += mat.B[3 * r] * e1[0] + mat.B[3 * r + 1] * e1[1] + mat.B[3 * r + 2] * e1[2];
            M[r] += mat.B[r] * e0[0] + mat.B[3 + r] * e0[1] + mat.B[6 + r] * e0[2];
        }
    }
    const double hw = 0.5 * w[g];
    em[g] = hw * (e0[0] * N[0] + e0[1] * N[1] + e0[2] * N[2]);
    eb[g] = hw * (e1[0] * M[0] + e1[1] * M[1] + e1[2] * M[2]);
}

// Export columns per face GP (Exporter::writeBinaryStep, plate/Exporter.h:34-91):
// X, Y, u, v, w, rotX, rotY (degrees) and the requested fields (codes 0..5:
// E_XX, E_YY, E_XY, K_XX, K_YY, K_XY; anything else writes 0). One wave per GP.
__global__ __launch_bounds__(EN_THREADS) void k_export_fields(int ng, const int* __restrict__ off,
                                                             const int* __restrict__ ids,
                                                             const double* __restrict__ sf, int64_t npairs,
                                                             const double* __restrict__ gx,
                                                             const double* __restrict__ gy,
                                                             const double* __restrict__ xu, int nf,
                                                             const int* __restrict__ fcode,
                                                             double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (EN_THREADS / 64) + (threadIdx.x >> 6);
    if (g >= ng) return;   // uniform per wave
    double p[18];
    for (int k = 0; k < 18; ++k) p[k] = 0.0;
    for (int k = off[g] + lane; k < off[g + 1]; k += 64) {
        const int v = ids[k];
        const double u0 = xu[3 * v], u1 = xu[3 * v + 1], u2 = xu[3 * v + 2];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double d = sf[q * npairs + k];
            p[3 * q] += u0 * d; p[3 * q + 1] += u1 * d; p[3 * q + 2] += u2 * d;
        }
    }
#pragma unroll
    for (int k = 0; k < 18; ++k) p[k] = wave_sum(p[k]);
    if (lane != 0) return;
    const double a1[3] = {1.0 + p[3], p[4], p[5]}, a2[3] = {p[6], 1.0 + p[7], p[8]};
    const double* rxx = p + 9; const double* ryy = p + 12; const double* rxy = p + 15;
    double n[3];
    cross3(a1, a2, n);
    const double J = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double b[3] = {n[0] / J, n[1] / J, n[2] / J};
    const double r2d = 180.0 / M_PI;
    double* o = out + static_cast<int64_t>(g) * (7 + nf);
    o[0] = gx[g]; o[1] = gy[g];
    o[2] = p[0]; o[3] = p[1]; o[4] = p[2];
    o[5] = atan2(-b[1], b[2]) * r2d;
    o[6] = atan2(b[0], b[2]) * r2d;
    for (int f = 0; f < nf; ++f) {
        double v = 0.0;
        switch (fcode[f]) {
            case 0: v = 0.5 * ((a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]) - 1.0); break;
            case 1: v = 0.5 * ((a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]) - 1.0); break;
            case 2: v = 0.5 * (a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2]); break;
            case 3: v = -(rxx[0] * b[0] + rxx[1] * b[1] + rxx[2] * b[2]); break;
            case 4: v = -(ryy[0] * b[0] + ryy[1] * b[1] + ryy[2] * b[2]); break;
            case 5: v = -(rxy[0] * b[0] + rxy[1] * b[1] + rxy[2] * b[2]); break;
            default: break;
        }
        o[7 + f] = v;
    }
}

struct CreaseSide {
    const int* off;
    const int* ids;
    const double* sf;
    int64_t npairs;
};

// r0,X and r0,Y of one crease GP side (Kinematics.h:66-91)
__device__ inline void side_gradient(const CreaseSide& s, int i, const double* __restrict__ xu, double* F) {
    double a1[3] = {1.0, 0.0, 0.0}, a2[3] = {0.0, 1.0, 0.0};
    for (int k = s.off[i]; k < s.off[i + 1]; ++k) {
        const int v = s.ids[k];
        const double px = s.sf[s.npairs + k], py = s.sf[2 * s.npairs + k];
#pragma unroll
        for (int d = 0; d < 3; ++d) { a1[d] += xu[3 * v + d] * px; a2[d] += xu[3 * v + d] * py; }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) { F[d] = a1[d]; F[3 + d] = a2[d]; }
}

__global__ void k_crease_angle(int n, CreaseSide L, CreaseSide R, const double* __restrict__ tx,
                               const double* __restrict__ ty, const double* __restrict__ w,
                               const double* __restrict__ xu, const double* __restrict__ theta0, double Ang0,
                               double Kt, const char* __restrict__ active, double* __restrict__ theta,
                               double* __restrict__ eh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double F[12];
    side_gradient(L, i, xu, F);
    side_gradient(R, i, xu, F + 6);
    const double th = fold_angle(F, tx[i], ty[i]);
    if (theta) theta[i] = th;
    if (eh) {
        const double d = th - (theta0 ? theta0[i] : Ang0);
        eh[i] = (active && !active[i]) ? 0.0 : w[i] * (0.5 * Kt * (d * d));
    }
}

// out[b] = sum of v_b[0 .. n_b): each thread sums a strided slice in index
// order, then a fixed pairwise tree in LDS.
__global__ __launch_bounds__(SUM_THREADS) void k_sum3(const double* __restrict__ v0, int64_t n0,
                                                       const double* __restrict__ v1, int64_t n1,
                                                       const double* __restrict__ v2, int64_t n2,
                                                       double* __restrict__ out) {
    __shared__ double red[SUM_THREADS];
    const double* v = blockIdx.x == 0 ? v0 : (blockIdx.x == 1 ? v1 : v2);
    const int64_t n = blockIdx.x == 0 ? n0 : (blockIdx.x == 1 ? n1 : n2);
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += SUM_THREADS) acc += v[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = SUM_THREADS / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

static CreaseSide crease_side(Domain& d) {
    return CreaseSide{d.off.p, d.ids.p, d.sf.p, d.npairs};
}

static void need_sf(const Domain& d, const char* what) {
    if (!d.have_sf || d.npairs < 0) throw Error(CVM_E_STATE, std::string(what) + ": shape functions not computed");
}

void launch_energies(Context& c, const double* xd, double* out_dev) {
    Domain* face = nullptr;
    for (Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) { face = d; break; }
    if (!face) throw Error(CVM_E_STATE, "no face domain");
    need_sf(*face, "energies");
    int64_t ncr = 0;
    for (const Fold& f : c.folds) ncr += c.doms[f.left]->ng;
    const int64_t ng = face->ng;
    c.en_scr.ensure(static_cast<size_t>(std::max<int64_t>(2 * ng + ncr, 1)));
    c.en_active.ensure(static_cast<size_t>(std::max<int64_t>(ncr, 1)));
    double* em = c.en_scr.p;
    double* eb = em + ng;
    double* eh = eb + ng;
    hipStream_t st = c.stream;
    MatC m;
    std::copy(c.A, c.A + 9, m.A); std::copy(c.B, c.B + 9, m.B); std::copy(c.D, c.D + 9, m.D);
    bool has_b = false;
    for (double v : c.B) has_b = has_b || v != 0.0;
    if (ng > 0) {
        const unsigned grid = static_cast<unsigned>((ng + EN_THREADS / 64 - 1) / (EN_THREADS / 64));
        auto kern = has_b ? k_face_energy<true> : k_face_energy<false>;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(EN_THREADS), 0, st, static_cast<int>(ng), face->off.p,
                           face->ids.p, face->sf.p, static_cast<int64_t>(face->npairs), face->d_w.p, xd,
                           natural_kappa(c, true), m, em, eb);
        CVM_HIP(hipGetLastError());
    }
    int64_t o = 0;
    for (Fold& f : c.folds) {
        Domain& L = *c.doms[f.left];
        Domain& R = *c.doms[f.right];
        need_sf(L, "hinge energy");
        need_sf(R, "hinge energy");
        const int n = L.ng;
        if (n == 0) continue;
        char* act = nullptr;
        if (!f.active.empty()) {
            CVM_HIP(hipMemcpyAsync(c.en_active.p + o, f.active.data(), n, hipMemcpyHostToDevice, st));
            act = c.en_active.p + o;
        }
        hipLaunchKernelGGL(k_crease_angle, dim3((n + 63) / 64), dim3(64), 0, st, n, crease_side(L), crease_side(R),
                           L.d_tx.p, L.d_ty.p, L.d_w.p, xd, f.theta0.empty() ? nullptr : f.d_theta0.p, f.Ang0, f.Kt,
                           act, nullptr, eh + o);
        CVM_HIP(hipGetLastError());
        o += n;
    }
    hipLaunchKernelGGL(k_sum3, dim3(3), dim3(SUM_THREADS), 0, st, em, ng, eb, ng, eh, ncr, out_dev);
    CVM_HIP(hipGetLastError());
}

void launch_export_fields(Context& c, const double* xd, int nf, const int* fcode_dev, double* out_dev) {
    Domain* face = nullptr;
    for (Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) { face = d; break; }
    if (!face) throw Error(CVM_E_STATE, "no face domain");
    need_sf(*face, "export fields");
    if (face->ng == 0) return;
    const unsigned grid = static_cast<unsigned>((face->ng + EN_THREADS / 64 - 1) / (EN_THREADS / 64));
    hipLaunchKernelGGL(k_export_fields, dim3(grid), dim3(EN_THREADS), 0, c.stream, face->ng, face->off.p,
                       face->ids.p, face->sf.p, static_cast<int64_t>(face->npairs), face->gx.p, face->gy.p, xd, nf,
                       fcode_dev, out_dev);
    CVM_HIP(hipGetLastError());
}

void launch_fold_angles(Context& c, int fold_id, const double* xd, double* theta_dev) {
    Fold& f = c.folds.at(static_cast<size_t>(fold_id));
    Domain& L = *c.doms[f.left];
    Domain& R = *c.doms[f.right];
    need_sf(L, "fold angles");
    need_sf(R, "fold angles");
    const int n = L.ng;
    if (n == 0) return;
    hipLaunchKernelGGL(k_crease_angle, dim3((n + 63) / 64), dim3(64), 0, c.stream, n, crease_side(L), crease_side(R),
                       L.d_tx.p, L.d_ty.p, L.d_w.p, xd, nullptr, 0.0, 0.0, nullptr, theta_dev, nullptr);
    CVM_HIP(hipGetLastError());
}

}  // namespace cvm
