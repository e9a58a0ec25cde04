// RKPM shape functions up to second derivatives on gfx950 (FP64).
//
// Reference algorithm: core/include/curvami/GlobalRKPM.h:19-397. Same
// quantities, reorganised for the device:
//  * one Gauss point per lane; the neighbour loop accumulates the six
//    window-weighted monomial moments mu_j[a,b] = sum_I phi_I^(j) xi^a eta^b
//    (j = phi, phi_x, phi_y, phi_xx, phi_yy, phi_xy; a + b <= 2p) in
//    registers. Every entry of the reference's moment matrices M, M_x, ...,
//    M_xy (GlobalRKPM.h:161-189) is a fixed combination of these
//    (product rule on x^a y^b), so the 6 x 21 packed entries (order 2) cost
//    90 accumulators and ~110 flop per pair instead of ~1.6 kflop;
//  * instead of the explicit 3x3 inverse + Schur bordering of the full
//    inverse and its derivatives (:203-345), M = L D L^T is factored once and
//    only the first column c = M^-1 e0 and its derivatives are solved:
//    c_x = -M^-1 M_x c, c_xx = -M^-1 (M_xx c + 2 M_x c_x), ... (row 0 of M^-1
//    is all the final assembly :363-397 reads). The LDL^T pivots D_k (k >= 3)
//    are exactly the reference's Schur complements q, so its pivot-collapse
//    test (:303-305) is applied to them; the 3x3 leading block is inverted by
//    its adjugate and checked with the reference's residual test (:214-220).
//  * the crease visibility (:86-112) uses per-GP indicator values and per-node
//    side signs computed once (the reference recomputes both per pair);
//  * two kernels: k_rkpm_coef (one lane per GP: moments, tests, solves) writes
//    c and its derivatives per GP to HBM, and k_rkpm_sf (lanes over the pairs of
//    a block of GPs) evaluates Psi and its derivatives (:363-397) with coalesced
//    stores at a much higher occupancy than the register-heavy first kernel.
//    Both are launched once for all domains of a call.
#include <algorithm>
#include <cstring>

#include "cvm_internal.h"

namespace cvm {

template <int P>
struct RK {
    static constexpr int NP = (P + 1) * (P + 2) / 2;
    static constexpr int DM = 2 * P;                              // max monomial degree
    static constexpr int NMON = (DM + 1) * (DM + 2) / 2;
    static constexpr int NCJ = 6 * NP;                            // c and its 5 derivatives
};

// Basis term k (degree ascending, then p_y ascending, GlobalRKPM.h:137-150)
// for orders up to 3: exponents of x and y.
__constant__ constexpr int kBEX[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0};
__constant__ constexpr int kBEY[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
template <int K> struct BE { static constexpr int x = kBEX[K], y = kBEY[K]; };
__host__ __device__ constexpr int midx(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }

template <int KER>
__device__ __forceinline__ void window(double s, double& w, double& dw, double& d2w) {
    if (KER == CVM_KERNEL_CUBIC_BSPLINE) {
        if (s >= 1.0) { w = 0.0; dw = 0.0; d2w = 0.0; return; }
        if (s <= 0.5) {
            w = 2.0 / 3.0 - 4.0 * s * s + 4.0 * s * s * s;
            dw = (s < 1e-14) ? 0.0 : -8.0 * s + 12.0 * s * s;
            d2w = -8.0 + 24.0 * s;
        } else {
            w = 4.0 / 3.0 - 4.0 * s + 4.0 * s * s - 4.0 / 3.0 * s * s * s;
            dw = -4.0 + 8.0 * s - 4.0 * s * s;
            d2w = 8.0 - 8.0 * s;
        }
        return;
    }
    // quartic spline, GlobalRKPM.h:15-17
    if (s >= 1.0) { w = 0.0; dw = 0.0; d2w = 0.0; return; }
    const double s2 = s * s;
    w = 1.0 - 6.0 * s2 + 8.0 * s2 * s - 3.0 * s2 * s2;
    dw = (s < 1e-14) ? 0.0 : -12.0 * s + 24.0 * s * s - 12.0 * s * s * s;
    d2w = -12.0 + 48.0 * s - 36.0 * s * s;
}

struct Jet { double v, x, y, xx, yy, xy; };

// Window jet of one pair without creases (GlobalRKPM.h:59-84).
template <int KER>
__device__ __forceinline__ Jet window_jet(double xi, double eta, double si) {
    double w, d, dd;
    window<KER>(si, w, d, dd);
    // one reciprocal per pair instead of five divisions
    const double is = (si == 0.0) ? 0.0 : 1.0 / si;
    const double is3 = is * is * is;
    const double dx = xi * is;
    const double dy = eta * is;
    // xi^2/s^3 and eta^2/s^3 as in GlobalRKPM.h:79-80
    const double d2x = (si == 0.0) ? dd : (xi * xi) * is3;
    const double d2y = (si == 0.0) ? dd : (eta * eta) * is3;
    Jet j;
    j.v = w; j.x = d * dx; j.y = d * dy;
    j.xx = dd * dx * dx + d * d2x;
    j.yy = dd * dy * dy + d * d2y;
    j.xy = (si == 0.0) ? 0.0 : dd * dx * dy + d * (-(xi * eta) * is3);
    return j;
}

// One crease's visibility factor W = (1 + sn f) / 2 applied to the jet
// (GlobalRKPM.h:86-112): sn = the node's side sign, f and its derivatives the
// GP's indicator values (in 1/rho units).
__device__ __forceinline__ void crease_cut(Jet& j, double sn, double f, double fx, double fy, double fxx,
                                           double fyy, double fxy, double rho) {
    const double W = 0.5 * (1.0 + sn * f);
    const double hs = 0.5 * sn;
    const double dWx = hs * fx * rho, dWy = hs * fy * rho;
    const double ddWxx = hs * fxx * rho * rho;
    const double ddWyy = hs * fyy * rho * rho;
    const double ddWxy = hs * fxy * rho * rho;
    const Jet o = j;
    j.v = o.v * W;
    j.x = o.x * W + o.v * dWx;
    j.y = o.y * W + o.v * dWy;
    j.xx = o.xx * W + 2.0 * o.x * dWx + o.v * ddWxx;
    j.yy = o.yy * W + 2.0 * o.y * dWy + o.v * ddWyy;
    j.xy = o.xy * W + o.x * dWy + o.y * dWx + o.v * ddWxy;
}

// Window jet of one pair with crease visibility (GlobalRKPM.h:59-112).
template <int KER>
__device__ __forceinline__ Jet pair_jet(double xi, double eta, double si, int I, int g, int ng, int nn,
                                        const double* __restrict__ ind, const int8_t* __restrict__ nsign,
                                        int nc, double rho) {
    Jet j = window_jet<KER>(xi, eta, si);
    for (int c = 0; c < nc; ++c) {
        const double sn = static_cast<double>(nsign[static_cast<size_t>(c) * nn + I]);
        const double* ig = ind + static_cast<size_t>(c) * 6 * ng + g;
        crease_cut(j, sn, ig[0], ig[ng], ig[2 * ng], ig[3 * ng], ig[4 * ng], ig[5 * ng], rho);
    }
    return j;
}

// Compile-time loop helper: f(Int<0>), ..., f(Int<N-1>).
template <int V> struct Int { static constexpr int value = V; };
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(Int<B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int RKPM_THREADS = RKPM_BLOCK;   // coefficient kernel: one lane per GP
constexpr int RKSF_THREADS = 512;          // shape-function kernel: lanes over a block's pairs

// One domain of a grouped launch: the domains of a call share one coefficient
// launch and one shape-function launch; block b belongs to the domain with the
// largest blk0 <= b (domains without GPs are not listed).
struct RkDom {
    const double* gx; const double* gy; const int* off;
    const int* il_row; const int* il_ids; const double* il_s;
    const int* ids; const double* s; const uint8_t* lgp;
    const double* ind; const int* gsheet; const SheetDev* sheets;
    double* sf; double* cj; int* status;
    int64_t npairs; int ng; int blk0;
};

__device__ __forceinline__ const RkDom& rk_domain(const RkDom* __restrict__ doms, int nd, int b) {
    int lo = 0, hi = nd;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (doms[mid].blk0 <= b) lo = mid; else hi = mid;
    }
    return doms[lo];
}

// Kernel 1: per GP, the moments, the singularity tests and c = M^-1 e0 with
// its five derivatives, written as a term-major record cj[g][t][6] =
// (c_t, cx_t, cy_t, cxx_t, cyy_t, cxy_t) for the shape-function kernel.
// Orders 1-2: at least 2 waves per SIMD (256 VGPRs). Order 3: one wave per
// SIMD, whose 512 registers (256 VGPRs + 256 AGPRs) hold the per-GP solve's
// working set that would otherwise spill to scratch memory.
template <int P, int KER, bool SPLIT>
__global__ __launch_bounds__(RKPM_THREADS) __attribute__((amdgpu_waves_per_eu(P >= 3 ? 1 : 2, P >= 3 ? 1 : 8)))
void k_rkpm_coef(const double* __restrict__ X, const double* __restrict__ Y, int nn,
                 const int8_t* __restrict__ nsign, const RkDom* __restrict__ doms, int nd) {
    using R = RK<P>;
    constexpr int NP = R::NP, NMON = R::NMON, DM = R::DM, NCJ = R::NCJ;
    const RkDom& D = rk_domain(doms, nd, blockIdx.x);
    const int lb = blockIdx.x - D.blk0;
    const double* __restrict__ gx = D.gx;
    const double* __restrict__ gy = D.gy;
    const int* __restrict__ off = D.off;
    const double* __restrict__ ind = D.ind;
    const int* __restrict__ gsheet = D.gsheet;
    const SheetDev* __restrict__ sheets = D.sheets;
    double* __restrict__ cjg = D.cj;
    int* __restrict__ status = D.status;       // n1, first1, n2, first2
    const int ng = D.ng;
    const int g = lb * RKPM_THREADS + threadIdx.x;
    if (g < ng) {
        const SheetDev& S = sheets[gsheet ? gsheet[g] : 0];
        const double rho = S.rho, irho = S.irho;
        const int nc = S.nc;
        const double px = gx[g], py = gy[g];
        const int cnt = off[g + 1] - off[g];
        // this lane's pairs in the interleaved lists: stride RKPM_THREADS
        const int* __restrict__ lid = D.il_ids + static_cast<size_t>(D.il_row[lb]) * RKPM_THREADS + threadIdx.x;
        const double* __restrict__ ls = D.il_s + static_cast<size_t>(D.il_row[lb]) * RKPM_THREADS + threadIdx.x;
        // One pass over the GP's pairs: f(window jet, xi, eta) per pair. Software
        // pipeline: the node and distance of pair t + 1 are loaded a pass ahead.
        auto for_pairs = [&](auto&& f) {
            // pair t + 1: node, distance and coordinates; pair t + 2: node, distance
            int In = cnt > 0 ? lid[0] : 0;
            double sn = cnt > 0 ? ls[0] : 0.0;
            double Xn = cnt > 0 ? X[In] : 0.0, Yn = cnt > 0 ? Y[In] : 0.0;
            int In2 = cnt > 1 ? lid[RKPM_THREADS] : 0;
            double sn2 = cnt > 1 ? ls[RKPM_THREADS] : 0.0;
            for (int t = 0; t < cnt; ++t) {
                const int I = In;
                const double si = sn;
                const double XI = Xn, YI = Yn;
                if (t + 1 < cnt) {
                    In = In2; sn = sn2;
                    Xn = X[In]; Yn = Y[In];
                    if (t + 2 < cnt) {
                        In2 = lid[static_cast<size_t>(t + 2) * RKPM_THREADS];
                        sn2 = ls[static_cast<size_t>(t + 2) * RKPM_THREADS];
                    }
                }
                const double xi = (px - XI) * irho;
                const double eta = (py - YI) * irho;
                f(pair_jet<KER>(xi, eta, si, I, g, ng, nn, ind, nsign, nc, rho), xi, eta);
            }
        };
        // ---- moments of monomials ---------------------------------------------
        // mu_j for j = phi, phi_x, phi_y and (not SPLIT) phi_xx, phi_yy, phi_xy.
        // SPLIT: the second-derivative moments enter only as M_xx c, M_yy c,
        // M_xy c, accumulated contracted in a second pass once c is known
        // (3 NP instead of 3 NMON registers).
        constexpr int NJ = SPLIT ? 3 : 6;
        double mu[NJ][NMON];
        static_for<0, NMON>([&](auto m) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) mu[j][m.value] = 0.0;
        });
        for_pairs([&](const Jet& jt, double xi, double eta) {
            double xp[DM + 1], yp[DM + 1];
            xp[0] = 1.0; yp[0] = 1.0;
#pragma unroll
            for (int q = 1; q <= DM; ++q) { xp[q] = xp[q - 1] * xi; yp[q] = yp[q - 1] * eta; }
#pragma unroll
            for (int d = 0; d <= DM; ++d)
#pragma unroll
                for (int bb = 0; bb <= d; ++bb) {
                    const int m = midx(d - bb, bb);
                    const double mon = xp[d - bb] * yp[bb];
                    mu[0][m] += jt.v * mon;
                    mu[1][m] += jt.x * mon;
                    mu[2][m] += jt.y * mon;
                    if constexpr (!SPLIT) {
                        mu[NJ - 3][m] += jt.xx * mon;
                        mu[NJ - 2][m] += jt.yy * mon;
                        mu[NJ - 1][m] += jt.xy * mon;
                    }
                }
        });
        // Entries of M and its derivatives from the monomial moments (product rule
        // on x^a y^b), all indices compile-time constants.
        auto MU = [&](int j, int a, int c) -> double {
            return (a < 0 || c < 0 || j >= NJ) ? 0.0 : mu[j][midx(a, c)];
        };
        auto Mv = [&](auto I, auto J) { return MU(0, BE<I.value>::x + BE<J.value>::x, BE<I.value>::y + BE<J.value>::y); };
        auto Mxv = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return MU(1, a, c) + a * MU(0, a - 1, c);
        };
        auto Myv = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return MU(2, a, c) + c * MU(0, a, c - 1);
        };
        auto Mxxv = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return MU(3, a, c) + 2 * a * MU(1, a - 1, c) + a * (a - 1) * MU(0, a - 2, c);
        };
        auto Myyv = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return MU(4, a, c) + 2 * c * MU(2, a, c - 1) + c * (c - 1) * MU(0, a, c - 2);
        };
        auto Mxyv = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return MU(5, a, c) + a * MU(2, a - 1, c) + c * MU(1, a, c - 1) + a * c * MU(0, a - 1, c - 1);
        };
        // the same without the second-derivative moment (SPLIT)
        auto Mxx0v = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return 2 * a * MU(1, a - 1, c) + a * (a - 1) * MU(0, a - 2, c);
        };
        auto Myy0v = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return 2 * c * MU(2, a, c - 1) + c * (c - 1) * MU(0, a, c - 2);
        };
        auto Mxy0v = [&](auto I, auto J) {
            constexpr int a = BE<I.value>::x + BE<J.value>::x, c = BE<I.value>::y + BE<J.value>::y;
            return a * MU(2, a - 1, c) + c * MU(1, a, c - 1) + a * c * MU(0, a - 1, c - 1);
        };

        // ---- singularity tests (GlobalRKPM.h:214-220 and :303-305) --------------
        int flag = 0;
        {
            double m[3][3];
            static_for<0, 3>([&](auto r) { static_for<0, 3>([&](auto c) { m[r.value][c.value] = Mv(r, c); }); });
            auto cof = [&](int i, int j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
            };
            const double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
            const double det = c0 * m[0][0] + c1 * m[1][0] + c2 * m[2][0];
            const double id = 1.0 / det;
            double inv[3][3];
            inv[0][0] = c0 * id; inv[0][1] = c1 * id; inv[0][2] = c2 * id;
            inv[1][0] = cof(0, 1) * id; inv[1][1] = cof(1, 1) * id; inv[1][2] = cof(2, 1) * id;
            inv[2][0] = cof(0, 2) * id; inv[2][1] = cof(1, 2) * id; inv[2][2] = cof(2, 2) * id;
            bool finite = true;
            double rmax = 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    finite = finite && isfinite(inv[r][c]);
                    const double pr = m[r][0] * inv[0][c] + m[r][1] * inv[1][c] + m[r][2] * inv[2][c];
                    rmax = fmax(rmax, fabs(pr - (r == c ? 1.0 : 0.0)));
                }
            if (!finite || !(rmax <= 1e-8)) flag |= 1;
        }

        // ---- L D L^T of M, then c = M^-1 e0 and its derivatives -----------------
        double L[NP][NP], Dg[NP];
        static_for<0, NP>([&](auto K) {
            constexpr int k = K.value;
            double dk = Mv(K, K);
#pragma unroll
            for (int j = 0; j < k; ++j) dk -= L[k][j] * L[k][j] * Dg[j];
            Dg[k] = dk;
            if (k >= 3 && !(fabs(dk) > 1e-12 * fmax(fabs(Mv(K, K)), 1.0))) flag |= 2;
            const double idk = 1.0 / dk;
            static_for<k + 1, NP>([&](auto I) {
                constexpr int i = I.value;
                double v = Mv(I, K);
#pragma unroll
                for (int j = 0; j < k; ++j) v -= L[i][j] * L[k][j] * Dg[j];
                L[i][k] = v * idk;
            });
        });
        auto solve = [&](double (&r)[NP]) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int j = 0; j < i; ++j) r[i] -= L[i][j] * r[j];
#pragma unroll
            for (int i = 0; i < NP; ++i) r[i] /= Dg[i];
#pragma unroll
            for (int i = NP - 1; i >= 0; --i)
#pragma unroll
                for (int j = i + 1; j < NP; ++j) r[i] -= L[j][i] * r[j];
        };
        double c[NP], cx[NP], cy[NP], txx[NP], tyy[NP], txy[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) c[i] = (i == 0) ? 1.0 : 0.0;
        solve(c);
        // second-order moment products with c first, so that mu_3..5 die early
        static_for<0, NP>([&](auto I) {
            double sxx = 0.0, syy = 0.0, sxy = 0.0, sx = 0.0, sy = 0.0;
            static_for<0, NP>([&](auto J) {
                if constexpr (SPLIT) {
                    sxx += Mxx0v(I, J) * c[J.value];
                    syy += Myy0v(I, J) * c[J.value];
                    sxy += Mxy0v(I, J) * c[J.value];
                } else {
                    sxx += Mxxv(I, J) * c[J.value];
                    syy += Myyv(I, J) * c[J.value];
                    sxy += Mxyv(I, J) * c[J.value];
                }
                sx += Mxv(I, J) * c[J.value];
                sy += Myv(I, J) * c[J.value];
            });
            txx[I.value] = sxx; tyy[I.value] = syy; txy[I.value] = sxy;
            cx[I.value] = -sx; cy[I.value] = -sy;
        });
        solve(cx);
        solve(cy);
        static_for<0, NP>([&](auto I) {
            double sxx = 0.0, syy = 0.0, sxy = 0.0;
            static_for<0, NP>([&](auto J) {
                sxx += 2.0 * Mxv(I, J) * cx[J.value];
                syy += 2.0 * Myv(I, J) * cy[J.value];
                sxy += Mxv(I, J) * cy[J.value] + Myv(I, J) * cx[J.value];
            });
            if constexpr (SPLIT) {
                txx[I.value] += sxx; tyy[I.value] += syy; txy[I.value] += sxy;
            } else {
                txx[I.value] = -(txx[I.value] + sxx);
                tyy[I.value] = -(tyy[I.value] + syy);
                txy[I.value] = -(txy[I.value] + sxy);
            }
        });
        double2* cj = reinterpret_cast<double2*>(cjg + static_cast<size_t>(g) * NCJ);   // NCJ even
        if constexpr (SPLIT) {
            // c, c_x, c_y are final: out of the registers before the second pass
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                cj[3 * i] = make_double2(c[i], cx[i]);
                cjg[static_cast<size_t>(g) * NCJ + 6 * i + 2] = cy[i];
            }
            // (M_xx c)_I += sum_k phi_xx,k p_I(k) C(k) with C = sum_J c_J p_J, and
            // likewise yy, xy: the second-derivative moment terms
            for_pairs([&](const Jet& jt, double xi, double eta) {
                double xp[P + 1], yp[P + 1];
                xp[0] = 1.0; yp[0] = 1.0;
#pragma unroll
                for (int q = 1; q <= P; ++q) { xp[q] = xp[q - 1] * xi; yp[q] = yp[q - 1] * eta; }
                double C = 0.0;
                static_for<0, NP>([&](auto T) { C += c[T.value] * (xp[BE<T.value>::x] * yp[BE<T.value>::y]); });
                const double ax = jt.xx * C, ay = jt.yy * C, axy = jt.xy * C;
                static_for<0, NP>([&](auto T) {
                    const double pt = xp[BE<T.value>::x] * yp[BE<T.value>::y];
                    txx[T.value] += ax * pt;
                    tyy[T.value] += ay * pt;
                    txy[T.value] += axy * pt;
                });
            });
#pragma unroll
            for (int i = 0; i < NP; ++i) { txx[i] = -txx[i]; tyy[i] = -tyy[i]; txy[i] = -txy[i]; }
        }
        solve(txx);
        solve(tyy);
        solve(txy);
        if constexpr (SPLIT) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                cjg[static_cast<size_t>(g) * NCJ + 6 * i + 3] = txx[i];
                cj[3 * i + 2] = make_double2(tyy[i], txy[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                cj[3 * i] = make_double2(c[i], cx[i]);
                cj[3 * i + 1] = make_double2(cy[i], txx[i]);
                cj[3 * i + 2] = make_double2(tyy[i], txy[i]);
            }
        }
        if (flag) {
            if (flag & 1) { atomicAdd(&status[0], 1); atomicMin(&status[1], g); }
            if (flag & 2) { atomicAdd(&status[2], 1); atomicMin(&status[3], g); }
        }
    }
}

// Kernel 2: Psi, grad Psi, Hessian of Psi (GlobalRKPM.h:363-397), one lane per
// pair over the contiguous pair range of a block of RKPM_BLOCK GPs (coalesced
// stores); the block's GP records are staged in LDS.
template <int P, int KER>
__global__ __launch_bounds__(RKSF_THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_rkpm_sf(
    const double* __restrict__ X, const double* __restrict__ Y, int nn, const int8_t* __restrict__ nsign,
    const RkDom* __restrict__ doms, int nd) {
    using R = RK<P>;
    constexpr int NP = R::NP, NCJ = R::NCJ;
    const RkDom& D = rk_domain(doms, nd, blockIdx.x);
    const int lb = blockIdx.x - D.blk0;
    const double* __restrict__ gx = D.gx;
    const double* __restrict__ gy = D.gy;
    const int* __restrict__ off = D.off;
    const int* __restrict__ ids = D.ids;
    const double* __restrict__ sarr = D.s;
    const uint8_t* __restrict__ lgp = D.lgp;
    const double* __restrict__ ind = D.ind;
    const int* __restrict__ gsheet = D.gsheet;
    const SheetDev* __restrict__ sheets = D.sheets;
    const double* __restrict__ cjg = D.cj;
    double* __restrict__ sf = D.sf;
    const int64_t npairs = D.npairs;
    const int ng = D.ng;
    constexpr int NCS = 2;                      // creases whose GP indicator values are staged
    __shared__ int nc_s[RKPM_BLOCK];
    __shared__ double gp_s[5][RKPM_BLOCK];      // gx, gy, rho, 1/rho, 1/rho^2
    __shared__ double ind_s[NCS * 6][RKPM_BLOCK];
    __shared__ double2 cj_s[RKPM_BLOCK * NCJ / 2];
    const int g0 = lb * RKPM_BLOCK;
    const int nblk = min(RKPM_BLOCK, ng - g0);
    // ---- per-GP data of the block into LDS ------------------------------------
    for (int i = threadIdx.x; i < nblk; i += RKSF_THREADS) {
        const int g = g0 + i;
        const SheetDev& S = sheets[gsheet ? gsheet[g] : 0];
        gp_s[0][i] = gx[g]; gp_s[1][i] = gy[g];
        gp_s[2][i] = S.rho; gp_s[3][i] = S.irho; gp_s[4][i] = S.irho2;
        nc_s[i] = S.nc;
        for (int c = 0; c < min(S.nc, NCS); ++c)
#pragma unroll
            for (int q = 0; q < 6; ++q) ind_s[c * 6 + q][i] = ind[(static_cast<size_t>(c) * 6 + q) * ng + g];
    }
    // the block's records are contiguous: nblk * NCJ doubles from g0 * NCJ
    const double2* __restrict__ src = reinterpret_cast<const double2*>(cjg + static_cast<size_t>(g0) * NCJ);
    for (int i = threadIdx.x; i < nblk * NCJ / 2; i += RKSF_THREADS) cj_s[i] = src[i];
    __syncthreads();

    const int p0 = off[g0], p1 = off[g0 + nblk];
    int k = p0 + threadIdx.x;
    // node, distance and GP of the lane's next pair, loaded a pass ahead
    int In = k < p1 ? ids[k] : 0;
    double sN = k < p1 ? sarr[k] : 0.0;
    int Ln = k < p1 ? lgp[k] : 0;
    for (; k < p1; k += RKSF_THREADS) {
        const int I = In;
        const double si = sN;
        const int lo = Ln;                     // GP of pair k within the block
        if (k + RKSF_THREADS < p1) {
            In = ids[k + RKSF_THREADS]; sN = sarr[k + RKSF_THREADS]; Ln = lgp[k + RKSF_THREADS];
        }
        const double XI = X[I], YI = Y[I];
        const double rho = gp_s[2][lo], irho = gp_s[3][lo], ir2 = gp_s[4][lo];
        const int nc = nc_s[lo];
        const double2* cj = cj_s + lo * (NCJ / 2);
        const double xi = (gp_s[0][lo] - XI) * irho;
        const double eta = (gp_s[1][lo] - YI) * irho;
        Jet jt = window_jet<KER>(xi, eta, si);
        for (int c = 0; c < nc; ++c) {
            const double sn = static_cast<double>(nsign[static_cast<size_t>(c) * nn + I]);
            if (c < NCS) {
                crease_cut(jt, sn, ind_s[c * 6][lo], ind_s[c * 6 + 1][lo], ind_s[c * 6 + 2][lo],
                           ind_s[c * 6 + 3][lo], ind_s[c * 6 + 4][lo], ind_s[c * 6 + 5][lo], rho);
            } else {
                const double* ig = ind + static_cast<size_t>(c) * 6 * ng + g0 + lo;
                crease_cut(jt, sn, ig[0], ig[ng], ig[2 * ng], ig[3 * ng], ig[4 * ng], ig[5 * ng], rho);
            }
        }
        double xp[P + 1], yp[P + 1];
        xp[0] = 1.0; yp[0] = 1.0;
#pragma unroll
        for (int q = 1; q <= P; ++q) { xp[q] = xp[q - 1] * xi; yp[q] = yp[q - 1] * eta; }
        double C = 0, Cx = 0, Cy = 0, Cxx = 0, Cyy = 0, Cxy = 0;
        static_for<0, NP>([&](auto T) {
            constexpr int t = T.value, ax = BE<t>::x, ay = BE<t>::y;
            const double p = xp[ax] * yp[ay];
            const double pxv = ax > 0 ? ax * xp[ax > 0 ? ax - 1 : 0] * yp[ay] : 0.0;
            const double pyv = ay > 0 ? ay * xp[ax] * yp[ay > 0 ? ay - 1 : 0] : 0.0;
            const double pxx = ax > 1 ? ax * (ax - 1) * xp[ax > 1 ? ax - 2 : 0] * yp[ay] : 0.0;
            const double pyy = ay > 1 ? ay * (ay - 1) * xp[ax] * yp[ay > 1 ? ay - 2 : 0] : 0.0;
            const double pxy = (ax > 0 && ay > 0) ? ax * ay * xp[ax > 0 ? ax - 1 : 0] * yp[ay > 0 ? ay - 1 : 0] : 0.0;
            const double2 r0 = cj[3 * t], r1 = cj[3 * t + 1], r2 = cj[3 * t + 2];
            const double v = r0.x, vx = r0.y, vy = r1.x, vxx = r1.y, vyy = r2.x, vxy = r2.y;
            C += v * p;
            Cx += vx * p + v * pxv;
            Cy += vy * p + v * pyv;
            Cxx += vxx * p + 2.0 * vx * pxv + v * pxx;
            Cyy += vyy * p + 2.0 * vy * pyv + v * pyy;
            Cxy += vxy * p + vx * pyv + vy * pxv + v * pxy;
        });
        sf[k] = C * jt.v;
        sf[npairs + k] = (Cx * jt.v + C * jt.x) * irho;
        sf[2 * npairs + k] = (Cy * jt.v + C * jt.y) * irho;
        sf[3 * npairs + k] = (Cxx * jt.v + 2.0 * Cx * jt.x + C * jt.xx) * ir2;
        sf[4 * npairs + k] = (Cyy * jt.v + 2.0 * Cy * jt.y + C * jt.yy) * ir2;
        sf[5 * npairs + k] = (Cxy * jt.v + Cx * jt.y + Cy * jt.x + C * jt.xy) * ir2;
    }
}

// Status words (n1, first1, n2, first2) of n domains.
__global__ void k_status_init(int* s, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s[4 * i] = 0; s[4 * i + 1] = 0x7fffffff; s[4 * i + 2] = 0; s[4 * i + 3] = 0x7fffffff;
    }
}

// Second-derivative moments contracted in a second pass from this order on:
// 6 NMON accumulators (order 3: 168 doubles) do not fit the register file.
#ifndef RKPM_SPLIT
#define RKPM_SPLIT(P) ((P) >= 3)
#endif

template <int P, int KER>
static void run(Context& c, hipStream_t st, const RkDom* tab, int nd, int nblocks) {
    hipLaunchKernelGGL((k_rkpm_coef<P, KER, RKPM_SPLIT(P)>), dim3(nblocks), dim3(RKPM_THREADS), 0, st, c.X.p,
                       c.Y.p, c.nn, c.d_node_sign.p, tab, nd);
    CVM_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_rkpm_sf<P, KER>), dim3(nblocks), dim3(RKSF_THREADS), 0, st, c.X.p, c.Y.p, c.nn,
                       c.d_node_sign.p, tab, nd);
    CVM_HIP(hipGetLastError());
}

static void check_args(int order, int kernel) {
    if (order < 1 || order > 3) throw Error(CVM_E_INVALID, "order must be 1, 2 or 3");
    if (kernel != CVM_KERNEL_QUARTIC && kernel != CVM_KERNEL_CUBIC_BSPLINE)
        throw Error(CVM_E_INVALID, "unknown window kernel");
}

// Host checks and the shape-function buffer of a domain before it joins a
// launch (its indicator values are evaluated before, for the whole group).
static void prepare_rkpm(Context& c, Domain& d) {
    if (d.npairs < 0) throw Error(CVM_E_STATE, "neighbours not computed for this domain");
    // the kernels read ind[(lc * 6 + k) * ng + g] and sign[lc * nn + I] for
    // lc < its sheet's crease count: both buffers must hold that many creases
    const size_t mc = static_cast<size_t>(c.max_sheet_creases());
    for (const SheetDev& t : d.h_sheets)
        if (t.nc < 0 || static_cast<size_t>(t.nc) > mc || t.crease0 + t.nc > static_cast<int>(c.creases.size()))
            throw Error(CVM_E_STATE, "crease table of the domain is out of date");
    if (mc > 0 && (d.ind.n < mc * 6 * static_cast<size_t>(d.ng) || c.d_node_sign.n < mc * static_cast<size_t>(c.nn)))
        throw Error(CVM_E_STATE, "crease indicator buffers are smaller than the crease count");
    if (d.ng > 0 && d.il_row.n != static_cast<size_t>((d.ng + RKPM_THREADS - 1) / RKPM_THREADS + 1))
        throw Error(CVM_E_STATE, "interleaved neighbour lists missing for this domain");
    if (d.ng > 0 && d.pair_lgp.n < static_cast<size_t>(d.npairs))
        throw Error(CVM_E_STATE, "pair GP indices missing for this domain");
    d.sf.alloc(std::max<int64_t>(6 * d.npairs, 1));
}

// Enqueues the shape functions of a group of domains in one coefficient and one
// shape-function launch; domain k's status goes to st[4 k .. 4 k + 3]. `cache`
// selects the context's descriptor table (uploaded only when it changes).
// Uploads go on the context's stream, the kernels on kst (default: the
// context's stream) after them.
static void enqueue_group(Context& c, const std::vector<Domain*>& ds, int order, int kernel, int* st, int cache,
                          const char* stage, hipStream_t kst) {
    if (ds.empty()) return;
    for (Domain* d : ds)
        if (d->npairs < 0) throw Error(CVM_E_STATE, "neighbours not computed for this domain");
    launch_indicators(c, ds, kst);
    for (Domain* d : ds) prepare_rkpm(c, *d);
    const size_t ncj = 6 * static_cast<size_t>((order + 1) * (order + 2) / 2);   // RK<order>::NCJ
    std::vector<RkDom> tab;
    std::vector<size_t> cj_off;
    int64_t nblocks = 0;
    size_t cj_total = 0;
    for (size_t k = 0; k < ds.size(); ++k) {
        Domain& d = *ds[k];
        if (d.ng == 0) continue;
        RkDom r{};
        r.gx = d.gx.p; r.gy = d.gy.p; r.off = d.off.p;
        r.il_row = d.il_row.p; r.il_ids = d.il_ids.p; r.il_s = d.il_s.p;
        r.ids = d.ids.p; r.s = d.s.p; r.lgp = d.pair_lgp.p;
        r.ind = d.ind.p; r.gsheet = c.batched ? d.d_gsheet.p : nullptr;
        r.sheets = reinterpret_cast<const SheetDev*>(d.d_sheets.p);
        r.sf = d.sf.p; r.cj = nullptr; r.status = st + 4 * k;
        r.npairs = d.npairs; r.ng = d.ng; r.blk0 = static_cast<int>(nblocks);
        cj_off.push_back(cj_total);            // records of the domain in the shared workspace
        cj_total += static_cast<size_t>(d.ng) * ncj;
        nblocks += (d.ng + RKPM_BLOCK - 1) / RKPM_BLOCK;
        if (nblocks > INT32_MAX) throw Error(CVM_E_INVALID, "too many Gauss points for one shape-function launch");
        tab.push_back(r);
    }
    if (!tab.empty()) {
        c.rk_cj.ensure(std::max<size_t>(cj_total, 1));
        for (size_t k = 0; k < tab.size(); ++k) tab[k].cj = c.rk_cj.p + cj_off[k];
        auto& host = c.rk_tab_host[cache];
        auto& dev = c.rk_tab[cache];
        const size_t bytes = tab.size() * sizeof(RkDom);
        if (host.size() != bytes || std::memcmp(host.data(), tab.data(), bytes) != 0) {
            host.assign(reinterpret_cast<const unsigned char*>(tab.data()),
                        reinterpret_cast<const unsigned char*>(tab.data()) + bytes);
            dev.upload(host, c.stream);
        }
    }
    const hipStream_t ks = fork_to(c, kst);
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(256), 0, ks, st, static_cast<int>(ds.size()));
    CVM_HIP(hipGetLastError());
    if (tab.empty()) return;
    const RkDom* dtab = reinterpret_cast<const RkDom*>(c.rk_tab[cache].p);
    const int nd = static_cast<int>(tab.size()), nb = static_cast<int>(nblocks);
    const int slot = timing_begin(c, stage, ks);
    if (kernel == CVM_KERNEL_QUARTIC) {
        if (order == 1) run<1, CVM_KERNEL_QUARTIC>(c, ks, dtab, nd, nb);
        else if (order == 2) run<2, CVM_KERNEL_QUARTIC>(c, ks, dtab, nd, nb);
        else run<3, CVM_KERNEL_QUARTIC>(c, ks, dtab, nd, nb);
    } else {
        if (order == 1) run<1, CVM_KERNEL_CUBIC_BSPLINE>(c, ks, dtab, nd, nb);
        else if (order == 2) run<2, CVM_KERNEL_CUBIC_BSPLINE>(c, ks, dtab, nd, nb);
        else run<3, CVM_KERNEL_CUBIC_BSPLINE>(c, ks, dtab, nd, nb);
    }
    timing_end(c, slot, ks);
}

static void finish_rkpm(Domain& d, int order, int kernel, const int* hs, int* code, int* nbad, int* first) {
    d.order = order; d.kernel = kernel;
    *code = CVM_OK; *nbad = 0; *first = 0;
    if (hs[0] > 0) { *code = CVM_E_SINGULAR_MOMENT; *nbad = hs[0]; *first = hs[1]; }
    else if (hs[2] > 0) { *code = CVM_E_PIVOT_COLLAPSE; *nbad = hs[2]; *first = hs[3]; }
    d.have_sf = (*code == CVM_OK);
}

static const char* stage_of(const Domain& d) { return d.kind == CVM_DOMAIN_FACE ? "rkpm_face" : "rkpm_line"; }

void launch_rkpm(Context& c, Domain& d, double /*rho*/, int order, int kernel, int* code, int* nbad, int* first) {
    check_args(order, kernel);
    launch_node_signs(c);
    c.d_status.ensure(4);
    enqueue_group(c, {&d}, order, kernel, c.d_status.p, 1, stage_of(d), nullptr);
    int hs[4];
    CVM_HIP(hipMemcpyAsync(hs, c.d_status.p, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    finish_rkpm(d, order, kernel, hs, code, nbad, first);
}

// Every domain in one launch pair, the small (non-face) domains first so that
// their blocks are dispatched first and finish under the face domain's; one
// status read-back at the end. On failure: the first failing domain (in domain
// order) and its status.
RkpmBatch enqueue_rkpm_all(Context& c, int order, int kernel, hipStream_t kst) {
    check_args(order, kernel);
    launch_node_signs(c, kst);
    const size_t nd = c.doms.size();
    c.d_status.ensure(std::max<size_t>(4 * nd, 4));
    // launch order: non-face domains, then face domains; status slot k of the
    // launch order, mapped back to domain order after the read-back
    RkpmBatch b;
    b.order = order; b.kernel = kernel;
    std::vector<Domain*> group;
    for (int pass = 0; pass < 2; ++pass)
        for (size_t i = 0; i < nd; ++i)
            if ((c.doms[i]->kind == CVM_DOMAIN_FACE) == (pass == 1)) { group.push_back(c.doms[i]); b.order_of.push_back(i); }
    enqueue_group(c, group, order, kernel, c.d_status.p, 0, "shape_functions", kst);
    return b;
}

void read_rkpm_status(Context& c, int* hs) {
    const size_t nd = c.doms.size();
    if (nd) CVM_HIP(hipMemcpyAsync(hs, c.d_status.p, 4 * nd * sizeof(int), hipMemcpyDeviceToHost, c.stream));
}

void finish_rkpm_all(Context& c, const RkpmBatch& b, int* dom, int* code, int* nbad, int* first) {
    std::vector<int> hs(4 * std::max<size_t>(c.doms.size(), 1));
    read_rkpm_status(c, hs.data());
    CVM_HIP(hipStreamSynchronize(c.stream));
    finish_rkpm_all(c, b, hs.data(), dom, code, nbad, first);
}

void finish_rkpm_all(Context& c, const RkpmBatch& b, const int* hs, int* dom, int* code, int* nbad, int* first) {
    const size_t nd = c.doms.size();
    std::vector<int> slot_of(nd);
    for (size_t k = 0; k < b.order_of.size(); ++k) slot_of[b.order_of[k]] = static_cast<int>(k);
    *dom = -1; *code = CVM_OK; *nbad = 0; *first = 0;
    for (size_t i = 0; i < nd; ++i) {
        int cd, nb, fb;
        finish_rkpm(*c.doms[i], b.order, b.kernel, hs + 4 * slot_of[i], &cd, &nb, &fb);
        if (cd != CVM_OK && *dom < 0) { *dom = static_cast<int>(i); *code = cd; *nbad = nb; *first = fb; }
    }
}

void launch_rkpm_all(Context& c, int order, int kernel, int* dom, int* code, int* nbad, int* first) {
    const RkpmBatch b = enqueue_rkpm_all(c, order, kernel, nullptr);
    finish_rkpm_all(c, b, dom, code, nbad, first);
}

}  // namespace cvm
