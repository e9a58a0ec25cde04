// ============================================================================
// CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++17, FP64 restatement of Curvami's per-Gauss-point hot path, used
// to check the HIP implementation. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it; the product library never links it.
//
// Each function cites the reference file:line it restates (paths relative to
// the Curvami source tree). Where Eigen supplies the arithmetic, the textbook
// operation is restated (3x3 / 2x2 inverse by adjugate, dot, cross, norm) with
// Eigen's evaluation order where it is known; exact rounding of Eigen's SIMD
// reductions is not reproduced (differences are at round-off level).
//
// Built with -ffp-contract=off so that the FRNN distance test rounds exactly
// like the device code (which uses explicit non-fused operations there).
// ============================================================================
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

struct V2 { double x, y; };
struct V3 {
    double x, y, z;
};
static inline V3 v3(double a, double b, double c) { V3 r; r.x = a; r.y = b; r.z = c; return r; }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 scl(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
static inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// 3x3 row-major matrix
struct M3 { double m[3][3]; };
static inline M3 m3zero() { M3 r; std::memset(&r, 0, sizeof(r)); return r; }
static inline M3 mmul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
static inline M3 madd(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
    return r;
}
static inline M3 msub(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j];
    return r;
}
static inline M3 mscl(const M3& a, double s) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
    return r;
}
static inline M3 mT(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
static inline V3 mv(const M3& a, V3 v) {
    return v3(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
              a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}

// Adjugate inverse of a 3x3 matrix (the textbook cofactor formula, with the
// cyclic cofactor indexing Eigen uses for Matrix3d::inverse; called at
// core/include/curvami/GlobalRKPM.h:214).
static inline double cof3(const M3& a, int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}
static inline M3 inv3(const M3& a) {
    const double c0 = cof3(a, 0, 0), c1 = cof3(a, 1, 0), c2 = cof3(a, 2, 0);
    const double det = c0 * a.m[0][0] + c1 * a.m[1][0] + c2 * a.m[2][0];
    const double id = 1.0 / det;
    M3 r;
    r.m[0][0] = c0 * id; r.m[0][1] = c1 * id; r.m[0][2] = c2 * id;
    r.m[1][0] = cof3(a, 0, 1) * id; r.m[1][1] = cof3(a, 1, 1) * id; r.m[1][2] = cof3(a, 2, 1) * id;
    r.m[2][0] = cof3(a, 0, 2) * id; r.m[2][1] = cof3(a, 1, 2) * id; r.m[2][2] = cof3(a, 2, 2) * id;
    return r;
}

// ---------------------------------------------------------------------------
// Crease indicator: core/include/curvami/Enrichment.h:10-92
// ---------------------------------------------------------------------------
struct FoldLine {
    std::vector<V2> v;
    bool ext_s = false, ext_e = false, closed = false;
};
struct Ind { double f, d, gx, gy, hxx, hyy, hxy; bool valid; };

static Ind indicator(V2 P, const FoldLine& fold) {
    Ind out;
    out.d = std::numeric_limits<double>::infinity();
    // Enrichment.h:29 returns f uninitialised for a polyline with < 2 vertices;
    // here it is 0 (no cut).
    out.f = 0.0; out.gx = out.gy = out.hxx = out.hyy = out.hxy = 0.0; out.valid = false;
    const size_t N = fold.v.size();
    if (N < 2) return out;
    V2 Cb = {0, 0}; size_t ib = 0; double tb = 0.0; V2 nb = {0, 0};
    const double inf = std::numeric_limits<double>::infinity();
    for (size_t i = 0; i < N - 1; ++i) {                               // :36-61
        const V2 A = fold.v[i], B = fold.v[i + 1];
        const double vx = B.x - A.x, vy = B.y - A.y;
        const double vsq = vx * vx + vy * vy;
        if (vsq == 0) continue;
        const double wx = P.x - A.x, wy = P.y - A.y;
        const double t = (wx * vx + wy * vy) / vsq;
        const double tmin = (i == 0 && fold.ext_s) ? -inf : 0.0;
        const double tmax = (i == N - 2 && fold.ext_e) ? inf : 1.0;
        const double tc = std::max(tmin, std::min(tmax, t));
        const double Cx = A.x + tc * vx, Cy = A.y + tc * vy;
        const double ex = P.x - Cx, ey = P.y - Cy;
        const double dsq = ex * ex + ey * ey;
        if (dsq < out.d * out.d) {
            out.d = std::sqrt(dsq);
            Cb.x = Cx; Cb.y = Cy; ib = i; tb = tc;
            const double sv = std::sqrt(vsq);
            nb.x = -vy / sv; nb.y = vx / sv;
        }
    }
    out.valid = (out.d > 1e-12);                                      // :63
    if (!out.valid) return out;                                       // :64-67 (f = 0)
    const double ux = (P.x - Cb.x) / out.d, uy = (P.y - Cb.y) / out.d;
    const double dp = ux * nb.x + uy * nb.y;
    const bool sc = (ib == 0 && tb == 0.0 && !fold.ext_s && !fold.closed);      // :72
    const bool ec = (ib == N - 2 && tb == 1.0 && !fold.ext_e && !fold.closed);  // :73
    if (sc || ec) {                                                   // :76-85
        out.f = dp;
        out.gx = (nb.x - out.f * ux) / out.d;
        out.gy = (nb.y - out.f * uy) / out.d;
        const double d2 = out.d * out.d;
        out.hxx = (3 * out.f * ux * ux - 2 * nb.x * ux - out.f) / d2;
        out.hyy = (3 * out.f * uy * uy - 2 * nb.y * uy - out.f) / d2;
        out.hxy = (3 * out.f * ux * uy - (nb.x * uy + ux * nb.y)) / d2;
    } else {                                                          // :86-89
        out.f = (dp >= 0.0) ? 1.0 : -1.0;
    }
    return out;
}

// ---------------------------------------------------------------------------
// Window functions. Quartic spline: core/include/curvami/GlobalRKPM.h:15-17.
// Cubic B-spline (kernel = 1): an added option with no reference counterpart,
// written from the standard piecewise-cubic B-spline on s in [0, 1].
// ---------------------------------------------------------------------------
static inline double kw(int kernel, double s) {
    if (kernel == 1) {
        if (s >= 1.0) return 0.0;
        if (s <= 0.5) return 2.0 / 3.0 - 4.0 * s * s + 4.0 * s * s * s;
        return 4.0 / 3.0 - 4.0 * s + 4.0 * s * s - 4.0 / 3.0 * s * s * s;
    }
    if (s >= 1.0) return 0.0;
    const double s2 = s * s;
    return 1.0 - 6.0 * s2 + 8.0 * s2 * s - 3.0 * s2 * s2;
}
static inline double kdw(int kernel, double s) {
    if (kernel == 1) {
        if (s >= 1.0 || s < 1e-14) return 0.0;
        if (s <= 0.5) return -8.0 * s + 12.0 * s * s;
        return -4.0 + 8.0 * s - 4.0 * s * s;
    }
    if (s >= 1.0 || s < 1e-14) return 0.0;
    return -12.0 * s + 24.0 * s * s - 12.0 * s * s * s;
}
static inline double kd2w(int kernel, double s) {
    if (kernel == 1) {
        if (s >= 1.0) return 0.0;
        if (s <= 0.5) return -8.0 + 24.0 * s;
        return 8.0 - 8.0 * s;
    }
    if (s >= 1.0) return 0.0;
    return -12.0 + 48.0 * s - 36.0 * s * s;
}

// ---------------------------------------------------------------------------
// Fixed-radius neighbour search: core/include/curvami/FRNN.h:23-78
// ---------------------------------------------------------------------------
struct NbrList {
    std::vector<int> off;   // ng + 1
    std::vector<int> ids;   // in FRNN order
    std::vector<double> s;  // sqrt(dist2) / rho
};

static void frnn(const std::vector<V2>& nodes, const std::vector<V2>& gps, double rho, NbrList& out) {
    out.off.assign(gps.size() + 1, 0);
    out.ids.clear(); out.s.clear();
    if (nodes.empty() || gps.empty()) return;
    V2 mn = nodes[0], mx = nodes[0];
    for (const auto& n : nodes) {                                     // :30-33
        mn.x = std::min(mn.x, n.x); mn.y = std::min(mn.y, n.y);
        mx.x = std::max(mx.x, n.x); mx.y = std::max(mx.y, n.y);
    }
    const double cell = rho;
    const int cols = static_cast<int>(std::ceil((mx.x - mn.x) / cell)) + 1;   // :34
    auto key = [&](V2 p) {                                           // :12-16
        const int x = static_cast<int>(std::floor((p.x - mn.x) / cell));
        const int y = static_cast<int>(std::floor((p.y - mn.y) / cell));
        return y * cols + x;
    };
    // unordered_map<int, vector> keyed by y*cols+x (:41-42); a std::map gives
    // the same per-key insertion order (ascending node index).
    std::map<int, std::vector<int>> grid;
    for (size_t i = 0; i < nodes.size(); ++i) grid[key(nodes[i])].push_back(static_cast<int>(i));
    const double rho2 = rho * rho;
    for (size_t q = 0; q < gps.size(); ++q) {                         // :56-77
        const int qx = static_cast<int>(std::floor((gps[q].x - mn.x) / cell));
        const int qy = static_cast<int>(std::floor((gps[q].y - mn.y) / cell));
        for (int dx = -1; dx <= 1; ++dx) {
            for (int dy = -1; dy <= 1; ++dy) {
                auto it = grid.find((qy + dy) * cols + (qx + dx));
                if (it == grid.end()) continue;
                for (int ni : it->second) {
                    const double ex = nodes[ni].x - gps[q].x, ey = nodes[ni].y - gps[q].y;
                    const double d2 = ex * ex + ey * ey;
                    if (d2 <= rho2) {
                        out.ids.push_back(ni);
                        out.s.push_back(std::sqrt(d2) / rho);
                    }
                }
            }
        }
        out.off[q + 1] = static_cast<int>(out.ids.size());
    }
}

// ---------------------------------------------------------------------------
// RKPM shape functions: core/include/curvami/GlobalRKPM.h:19-397
// ---------------------------------------------------------------------------
struct Shape {
    std::vector<double> psi, dpx, dpy, dxx, dyy, dxy;  // per pair, FRNN order
};

struct RkpmStatus { int code; int nbad; int first; };

static RkpmStatus rkpm(const std::vector<V2>& nodes, const std::vector<V2>& gps, const NbrList& nl,
                       double rho, int order, int kernel, const std::vector<FoldLine>& creases,
                       Shape& out, int nthreads) {
    const size_t ng = gps.size();
    const size_t nnz = nl.ids.size();
    const int np = (order + 1) * (order + 2) / 2;                     // :27
    std::vector<int> K(np * np);                                      // :30-36
    int nm = 0;
    for (int i = 0; i < np; ++i) for (int j = i; j < np; ++j) { K[i * np + j] = K[j * np + i] = nm++; }

    out.psi.assign(nnz, 0); out.dpx.assign(nnz, 0); out.dpy.assign(nnz, 0);
    out.dxx.assign(nnz, 0); out.dyy.assign(nnz, 0); out.dxy.assign(nnz, 0);

    // Indicator values of each GP and node (a pure function of the point; the
    // reference re-evaluates them per pair at :91-92 with identical results).
    const size_t nc = creases.size();
    std::vector<Ind> indG(ng * nc);
    std::vector<double> snN(nodes.size() * nc);
    for (size_t c = 0; c < nc; ++c) {
        for (size_t g = 0; g < ng; ++g) indG[g * nc + c] = indicator(gps[g], creases[c]);
        for (size_t n = 0; n < nodes.size(); ++n) {
            const Ind iN = indicator(nodes[n], creases[c]);
            snN[n * nc + c] = (iN.f >= 0.0) ? 1.0 : -1.0;             // :93
        }
    }

    std::vector<char> bad1(ng, 0), bad2(ng, 0);
    // exponents of the basis: degree d ascending, then p_y ascending (:137-150)
    std::vector<int> ex(np), ey(np);
    {
        int idx = 0;
        for (int d = 0; d <= order; ++d)
            for (int py = 0; py <= d; ++py) { ex[idx] = d - py; ey[idx] = py; ++idx; }
    }

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long long gg = 0; gg < static_cast<long long>(ng); ++gg) {
        const size_t g = static_cast<size_t>(gg);
        const int b = nl.off[g], e = nl.off[g + 1];
        const int n = e - b;
        // Per-pair kernel jets (:59-84) and visibility (:86-112)
        std::vector<double> PSI(n), dPx(n), dPy(n), d2Pxx(n), d2Pyy(n), d2Pxy(n), CSI(n), ETA(n);
        for (int k = 0; k < n; ++k) {
            const int I = nl.ids[b + k];
            const double xi = (gps[g].x - nodes[I].x) / rho;
            const double eta = (gps[g].y - nodes[I].y) / rho;
            CSI[k] = xi; ETA[k] = eta;
            const double si = nl.s[b + k];
            const double wi = kw(kernel, si), d = kdw(kernel, si), dd = kd2w(kernel, si);
            const double dx = (si == 0.0) ? 0.0 : xi / si;
            const double dy = (si == 0.0) ? 0.0 : eta / si;
            PSI[k] = wi; dPx[k] = d * dx; dPy[k] = d * dy;
            const double si3 = si * si * si;
            // The reference uses xi^2/s^3 (not eta^2/s^3) here (:79-80);
            // restated as written.
            const double d2x = (si == 0.0) ? dd : (xi * xi) / si3;
            const double d2y = (si == 0.0) ? dd : (eta * eta) / si3;
            d2Pxx[k] = dd * dx * dx + d * d2x;
            d2Pyy[k] = dd * dy * dy + d * d2y;
            d2Pxy[k] = (si == 0.0) ? 0.0 : dd * dx * dy + d * (-(xi * eta) / si3);
            for (size_t c = 0; c < nc; ++c) {
                const Ind& iG = indG[g * nc + c];
                const double sn = snN[static_cast<size_t>(I) * nc + c];
                const double W = 0.5 * (1.0 + sn * iG.f);
                const double dWx = 0.5 * sn * iG.gx * rho, dWy = 0.5 * sn * iG.gy * rho;
                const double hs = 0.5 * sn;
                const double ddWxx = hs * iG.hxx * rho * rho;
                const double ddWyy = hs * iG.hyy * rho * rho;
                const double ddWxy = hs * iG.hxy * rho * rho;
                const double oW = PSI[k], px = dPx[k], py = dPy[k];
                const double pxx = d2Pxx[k], pyy = d2Pyy[k], pxy = d2Pxy[k];
                PSI[k] *= W;
                dPx[k] = px * W + oW * dWx;
                dPy[k] = py * W + oW * dWy;
                d2Pxx[k] = pxx * W + 2.0 * px * dWx + oW * ddWxx;
                d2Pyy[k] = pyy * W + 2.0 * py * dWy + oW * ddWyy;
                d2Pxy[k] = pxy * W + px * dWy + py * dWx + oW * ddWxy;
            }
        }
        // Basis values and derivatives (:122-152)
        std::vector<double> P(np * n, 0.0), PX(np * n, 0.0), PY(np * n, 0.0),
            PXX(np * n, 0.0), PYY(np * n, 0.0), PXY(np * n, 0.0);
        std::vector<double> xp(order + 1), yp(order + 1);
        for (int k = 0; k < n; ++k) {
            xp[0] = 1.0; yp[0] = 1.0;
            for (int p = 1; p <= order; ++p) { xp[p] = xp[p - 1] * CSI[k]; yp[p] = yp[p - 1] * ETA[k]; }
            for (int idx = 0; idx < np; ++idx) {
                const int px = ex[idx], py = ey[idx];
                P[idx * n + k] = xp[px] * yp[py];
                if (px > 0) PX[idx * n + k] = static_cast<double>(px) * xp[px - 1] * yp[py];
                if (py > 0) PY[idx * n + k] = static_cast<double>(py) * xp[px] * yp[py - 1];
                if (px > 1) PXX[idx * n + k] = static_cast<double>(px * (px - 1)) * xp[px - 2] * yp[py];
                if (py > 1) PYY[idx * n + k] = static_cast<double>(py * (py - 1)) * xp[px] * yp[py - 2];
                if (px > 0 && py > 0) PXY[idx * n + k] = static_cast<double>(px * py) * xp[px - 1] * yp[py - 1];
            }
        }
        // Moments (:161-189)
        std::vector<double> M(nm, 0.0), dMx(nm, 0.0), dMy(nm, 0.0), d2Mxx(nm, 0.0), d2Myy(nm, 0.0), d2Mxy(nm, 0.0);
        for (int k = 0; k < n; ++k) {
            for (int ip = 0; ip < np; ++ip) {
                for (int jp = ip; jp < np; ++jp) {
                    const int q = K[ip * np + jp];
                    const double pi = P[ip * n + k], pj = P[jp * n + k];
                    const double dpix = PX[ip * n + k], dpjx = PX[jp * n + k];
                    const double dpiy = PY[ip * n + k], dpjy = PY[jp * n + k];
                    M[q] += pi * pj * PSI[k];
                    dMx[q] += pi * pj * dPx[k] + (dpix * pj + pi * dpjx) * PSI[k];
                    dMy[q] += pi * pj * dPy[k] + (dpiy * pj + pi * dpjy) * PSI[k];
                    d2Mxx[q] += d2Pxx[k] * pi * pj + 2.0 * dPx[k] * (dpix * pj + pi * dpjx)
                                + PSI[k] * (PXX[ip * n + k] * pj + 2.0 * dpix * dpjx + pi * PXX[jp * n + k]);
                    d2Myy[q] += d2Pyy[k] * pi * pj + 2.0 * dPy[k] * (dpiy * pj + pi * dpjy)
                                + PSI[k] * (PYY[ip * n + k] * pj + 2.0 * dpiy * dpjy + pi * PYY[jp * n + k]);
                    d2Mxy[q] += d2Pxy[k] * pi * pj + dPx[k] * (dpiy * pj + pi * dpjy)
                                + dPy[k] * (dpix * pj + pi * dpjx)
                                + PSI[k] * (PXY[ip * n + k] * pj + dpix * dpjy + dpiy * dpjx + pi * PXY[jp * n + k]);
                }
            }
        }
        // 3x3 leading block inverse and its derivatives (:203-239)
        std::vector<double> Mi(nm, 0.0), Mix(nm, 0.0), Miy(nm, 0.0), Mixx(nm, 0.0), Miyy(nm, 0.0), Mixy(nm, 0.0);
        auto fill3 = [&](const std::vector<double>& S) {
            M3 r;
            for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) r.m[a][c] = S[K[a * np + c]];
            return r;
        };
        {
            const M3 m3 = fill3(M), dmx = fill3(dMx), dmy = fill3(dMy);
            const M3 dxx3 = fill3(d2Mxx), dyy3 = fill3(d2Myy), dxy3 = fill3(d2Mxy);
            const M3 inv = inv3(m3);
            bool finite = true;
            for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) finite = finite && std::isfinite(inv.m[a][c]);
            double rmax = 0.0;
            const M3 prod = mmul(m3, inv);
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c)
                    rmax = std::max(rmax, std::fabs(prod.m[a][c] - (a == c ? 1.0 : 0.0)));
            if (!finite || rmax > 1e-8 || std::isnan(rmax)) bad1[g] = 1;   // :217-220
            const M3 ninv = mscl(inv, -1.0);
            const M3 ix = mmul(mmul(ninv, dmx), inv);
            const M3 iy = mmul(mmul(ninv, dmy), inv);
            const M3 ixx = mmul(mmul(inv, msub(mmul(mmul(mscl(dmx, 2.0), inv), dmx), dxx3)), inv);
            const M3 iyy = mmul(mmul(inv, msub(mmul(mmul(mscl(dmy, 2.0), inv), dmy), dyy3)), inv);
            const M3 ixy = mmul(mmul(inv, msub(madd(mmul(mmul(dmx, inv), dmy), mmul(mmul(dmy, inv), dmx)), dxy3)), inv);
            for (int r = 0; r < 3; ++r)
                for (int c = r; c < 3; ++c) {
                    const int q = K[r * np + c];
                    Mi[q] = inv.m[r][c]; Mix[q] = ix.m[r][c]; Miy[q] = iy.m[r][c];
                    Mixx[q] = ixx.m[r][c]; Miyy[q] = iyy.m[r][c]; Mixy[q] = ixy.m[r][c];
                }
        }
        // Schur-complement bordering for the remaining basis terms (:259-345).
        // A GP flagged singular above still runs this (as in the reference);
        // the singular report takes precedence.
        for (int it = 3; it < np; ++it) {
            std::vector<double> bv(it), dbx(it), dby(it), d2bxx(it), d2byy(it), d2bxy(it);
            for (int j = 0; j < it; ++j) {
                const int q = K[j * np + it];
                bv[j] = M[q]; dbx[j] = dMx[q]; dby[j] = dMy[q];
                d2bxx[j] = d2Mxx[q]; d2byy[j] = d2Myy[q]; d2bxy[j] = d2Mxy[q];
            }
            auto matvec = [&](const std::vector<double>& A, const std::vector<double>& x) {
                std::vector<double> y(it, 0.0);
                for (int c = 0; c < it; ++c)
                    for (int r = 0; r < it; ++r) y[r] += A[K[r * np + c]] * x[c];
                return y;
            };
            auto vdot = [&](const std::vector<double>& a, const std::vector<double>& c) {
                double s = 0.0;
                for (int j = 0; j < it; ++j) s += a[j] * c[j];
                return s;
            };
            const std::vector<double> v = matvec(Mi, bv);
            std::vector<double> vx(it), vy(it), vxx(it), vyy(it), vxy(it);
            {
                const auto a1 = matvec(Mix, bv), a2 = matvec(Mi, dbx);
                const auto b1 = matvec(Miy, bv), b2 = matvec(Mi, dby);
                const auto c1 = matvec(Mixx, bv), c2 = matvec(Mix, dbx), c3 = matvec(Mi, d2bxx);
                const auto e1 = matvec(Miyy, bv), e2 = matvec(Miy, dby), e3 = matvec(Mi, d2byy);
                const auto f1 = matvec(Mixy, bv), f2 = matvec(Mix, dby), f3 = matvec(Miy, dbx), f4 = matvec(Mi, d2bxy);
                for (int j = 0; j < it; ++j) {
                    vx[j] = a1[j] + a2[j];
                    vy[j] = b1[j] + b2[j];
                    vxx[j] = c1[j] + 2.0 * c2[j] + c3[j];
                    vyy[j] = e1[j] + 2.0 * e2[j] + e3[j];
                    vxy[j] = f1[j] + f2[j] + f3[j] + f4[j];
                }
            }
            const int kii = K[it * np + it];
            const double q0 = M[kii] - vdot(bv, v);
            const double qx = dMx[kii] - (vdot(dbx, v) + vdot(bv, vx));
            const double qy = dMy[kii] - (vdot(dby, v) + vdot(bv, vy));
            const double qxx = d2Mxx[kii] - (vdot(d2bxx, v) + 2.0 * vdot(dbx, vx) + vdot(bv, vxx));
            const double qyy = d2Myy[kii] - (vdot(d2byy, v) + 2.0 * vdot(dby, vy) + vdot(bv, vyy));
            const double qxy = d2Mxy[kii] - (vdot(d2bxy, v) + vdot(dbx, vy) + vdot(dby, vx) + vdot(bv, vxy));
            if (!(std::fabs(q0) > 1e-12 * std::max(std::fabs(M[kii]), 1.0))) {   // :303-305
                bad2[g] = 1;
            }
            const double iQ = 1.0 / q0, iQ2 = iQ * iQ, iQ3 = iQ2 * iQ;
            for (int r = 0; r < it; ++r)
                for (int c = r; c < it; ++c) {
                    const int q = K[r * np + c];
                    Mi[q] += v[r] * v[c] * iQ;
                    Mix[q] += (vx[r] * v[c] + v[r] * vx[c]) * iQ - v[r] * v[c] * qx * iQ2;
                    Miy[q] += (vy[r] * v[c] + v[r] * vy[c]) * iQ - v[r] * v[c] * qy * iQ2;
                    Mixx[q] += (vxx[r] * v[c] + 2.0 * vx[r] * vx[c] + v[r] * vxx[c]) * iQ
                               - (vx[r] * v[c] + v[r] * vx[c]) * qx * iQ2 * 2.0
                               + v[r] * v[c] * (2.0 * qx * qx * iQ3 - qxx * iQ2);
                    Miyy[q] += (vyy[r] * v[c] + 2.0 * vy[r] * vy[c] + v[r] * vyy[c]) * iQ
                               - (vy[r] * v[c] + v[r] * vy[c]) * qy * iQ2 * 2.0
                               + v[r] * v[c] * (2.0 * qy * qy * iQ3 - qyy * iQ2);
                    Mixy[q] += (vxy[r] * v[c] + vx[r] * vy[c] + vy[r] * vx[c] + v[r] * vxy[c]) * iQ
                               - (vx[r] * v[c] + v[r] * vx[c]) * qy * iQ2
                               - (vy[r] * v[c] + v[r] * vy[c]) * qx * iQ2
                               + v[r] * v[c] * (2.0 * qx * qy * iQ3 - qxy * iQ2);
                }
            for (int j = 0; j < it; ++j) {
                const int q = K[j * np + it];
                Mi[q] = -v[j] * iQ;
                Mix[q] = -vx[j] * iQ + v[j] * qx * iQ2;
                Miy[q] = -vy[j] * iQ + v[j] * qy * iQ2;
                Mixx[q] = -vxx[j] * iQ + 2.0 * vx[j] * qx * iQ2 + v[j] * (qxx * iQ2 - 2.0 * qx * qx * iQ3);
                Miyy[q] = -vyy[j] * iQ + 2.0 * vy[j] * qy * iQ2 + v[j] * (qyy * iQ2 - 2.0 * qy * qy * iQ3);
                Mixy[q] = -vxy[j] * iQ + vx[j] * qy * iQ2 + vy[j] * qx * iQ2 + v[j] * (qxy * iQ2 - 2.0 * qx * qy * iQ3);
            }
            Mi[kii] = iQ;
            Mix[kii] = -qx * iQ2;
            Miy[kii] = -qy * iQ2;
            Mixx[kii] = 2.0 * qx * qx * iQ3 - qxx * iQ2;
            Miyy[kii] = 2.0 * qy * qy * iQ3 - qyy * iQ2;
            Mixy[kii] = 2.0 * qx * qy * iQ3 - qxy * iQ2;
        }
        // Psi, grad Psi, Hessian of Psi (:363-397)
        for (int k = 0; k < n; ++k) {
            double C = 0, dCx = 0, dCy = 0, d2Cxx = 0, d2Cyy = 0, d2Cxy = 0;
            for (int jp = 0; jp < np; ++jp) {
                const int q = K[0 * np + jp];
                const double v = Mi[q], vx = Mix[q], vy = Miy[q], vxx = Mixx[q], vyy = Miyy[q], vxy = Mixy[q];
                const double p = P[jp * n + k], px = PX[jp * n + k], py = PY[jp * n + k];
                C += v * p;
                dCx += vx * p + v * px;
                dCy += vy * p + v * py;
                d2Cxx += vxx * p + 2.0 * vx * px + v * PXX[jp * n + k];
                d2Cyy += vyy * p + 2.0 * vy * py + v * PYY[jp * n + k];
                d2Cxy += vxy * p + vx * py + vy * px + v * PXY[jp * n + k];
            }
            const size_t o = static_cast<size_t>(b + k);
            out.psi[o] = C * PSI[k];
            out.dpx[o] = (dCx * PSI[k] + C * dPx[k]) / rho;
            out.dpy[o] = (dCy * PSI[k] + C * dPy[k]) / rho;
            out.dxx[o] = (d2Cxx * PSI[k] + 2.0 * dCx * dPx[k] + C * d2Pxx[k]) / (rho * rho);
            out.dyy[o] = (d2Cyy * PSI[k] + 2.0 * dCy * dPy[k] + C * d2Pyy[k]) / (rho * rho);
            out.dxy[o] = (d2Cxy * PSI[k] + dCx * dPy[k] + dCy * dPx[k] + C * d2Pxy[k]) / (rho * rho);
        }
    }
    // Singular 3x3 blocks are reported first (:241-256); the pivot check only
    // runs when none is singular (:347-359).
    RkpmStatus st{0, 0, 0};
    size_t n1 = 0, f1 = 0, n2 = 0, f2 = 0;
    for (size_t g = 0; g < ng; ++g) {
        if (bad1[g]) { if (!n1) f1 = g; ++n1; }
        if (bad2[g]) { if (!n2) f2 = g; ++n2; }
    }
    if (n1) { st.code = 3; st.nbad = static_cast<int>(n1); st.first = static_cast<int>(f1); return st; }
    if (n2) { st.code = 4; st.nbad = static_cast<int>(n2); st.first = static_cast<int>(f2); return st; }
    return st;
}

// ---------------------------------------------------------------------------
// Laminate: plate/Constitutive.h:10-50 (computeABD), :57-70 (evaluate)
// ---------------------------------------------------------------------------
struct Ply { double E1, E2, nu12, G12, angle_deg, thickness; };
struct ABD { M3 A, B, D; };

static ABD computeABD(const std::vector<Ply>& plies) {
    ABD r; r.A = m3zero(); r.B = m3zero(); r.D = m3zero();
    double h = 0;
    for (const auto& p : plies) h += p.thickness;
    double z0 = -h / 2.0;
    for (const auto& p : plies) {
        const double nu21 = (p.E2 / p.E1) * p.nu12;
        const double den = 1.0 - p.nu12 * nu21;
        M3 Q = m3zero();
        Q.m[0][0] = p.E1 / den; Q.m[1][1] = p.E2 / den;
        Q.m[0][1] = Q.m[1][0] = (p.nu12 * p.E2) / den;
        Q.m[2][2] = p.G12;
        const double th = p.angle_deg * M_PI / 180.0;
        const double m = std::cos(th), n = std::sin(th);
        const double m2 = m * m, n2 = n * n, m4 = m2 * m2, n4 = n2 * n2;
        M3 Qb;
        Qb.m[0][0] = Q.m[0][0] * m4 + 2.0 * (Q.m[0][1] + 2.0 * Q.m[2][2]) * m2 * n2 + Q.m[1][1] * n4;
        Qb.m[1][1] = Q.m[0][0] * n4 + 2.0 * (Q.m[0][1] + 2.0 * Q.m[2][2]) * m2 * n2 + Q.m[1][1] * m4;
        Qb.m[0][1] = Qb.m[1][0] = (Q.m[0][0] + Q.m[1][1] - 4.0 * Q.m[2][2]) * m2 * n2 + Q.m[0][1] * (m4 + n4);
        Qb.m[0][2] = Qb.m[2][0] = (Q.m[0][0] - Q.m[0][1] - 2.0 * Q.m[2][2]) * n * m * m2
                                  + (Q.m[0][1] - Q.m[1][1] + 2.0 * Q.m[2][2]) * n2 * n * m;
        Qb.m[1][2] = Qb.m[2][1] = (Q.m[0][0] - Q.m[0][1] - 2.0 * Q.m[2][2]) * m * n * n2
                                  + (Q.m[0][1] - Q.m[1][1] + 2.0 * Q.m[2][2]) * m2 * m * n;
        Qb.m[2][2] = (Q.m[0][0] + Q.m[1][1] - 2.0 * Q.m[0][1] - 2.0 * Q.m[2][2]) * m2 * n2 + Q.m[2][2] * (m4 + n4);
        const double z1 = z0 + p.thickness;
        r.A = madd(r.A, mscl(Qb, z1 - z0));
        r.B = madd(r.B, mscl(mscl(Qb, 0.5), z1 * z1 - z0 * z0));
        r.D = madd(r.D, mscl(mscl(Qb, 1.0 / 3.0), z1 * z1 * z1 - z0 * z0 * z0));
        z0 = z1;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Natural curvature: plate/Kinematics.h:9-45. A band along Y centred at x0,
// kappa_bar(X) = sign * theta0 * g(X) eX eX^T with g(X) = exp(-((X - x0) / w)^2)
// / (sqrt(pi) w); it enters the elastic bending strain E1_el = E1 + kappa_bar
// (:114-115), which the resultants (Constitutive.h:64-65) and the bending
// energy (Meshless.h:212) use. Inactive by default.
// ---------------------------------------------------------------------------
struct NatCurv {
    bool active = false;
    double x0 = 0.0, width = 1.0, theta0 = 0.0, sign = -1.0;
    double g(double X) const {                                        // :25-28
        const double t = (X - x0) / width;
        return std::exp(-t * t) / (1.7724538509055160273 * width);
    }
    double kappa00(double X) const { return active ? sign * theta0 * g(X) : 0.0; }   // :29-33
};

// per GP of a domain: kappa_bar(0, 0) at the GP's reference position
static std::vector<double> natural_kappa(const NatCurv& nc, const std::vector<V2>& X) {
    std::vector<double> k(X.size(), 0.0);
    if (nc.active)
        for (size_t g = 0; g < X.size(); ++g) k[g] = nc.kappa00(X[g].x);
    return k;
}

// ---------------------------------------------------------------------------
// Kirchhoff-Love kinematics: plate/Kinematics.h:61-120 (the natural curvature
// is added to E1 by the callers that form E1_el)
// ---------------------------------------------------------------------------
struct Kin {
    V3 a1, a2;            // F0 columns (r0,X, r0,Y)
    V3 b;                 // unit normal
    double Ci[2][2];      // C0^{-1}
    double E0[2][2], E1[2][2];
    V3 rXX, rYY, rXY;
};

static Kin kinematics(const int* ids, int n, const double* dpx, const double* dpy,
                      const double* dxx, const double* dyy, const double* dxy, const double* u) {
    Kin k;
    V3 rX = v3(1, 0, 0), rY = v3(0, 1, 0), rXX = v3(0, 0, 0), rYY = v3(0, 0, 0), rXY = v3(0, 0, 0);
    for (int i = 0; i < n; ++i) {                                     // :75-88
        const double* ui = u + 3 * static_cast<size_t>(ids[i]);
        const V3 U = v3(ui[0], ui[1], ui[2]);
        rX = add(rX, scl(U, dpx[i]));
        rY = add(rY, scl(U, dpy[i]));
        rXX = add(rXX, scl(U, dxx[i]));
        rYY = add(rYY, scl(U, dyy[i]));
        rXY = add(rXY, scl(U, dxy[i]));
    }
    k.a1 = rX; k.a2 = rY;
    double C[2][2];
    C[0][0] = dot(rX, rX); C[1][1] = dot(rY, rY); C[0][1] = C[1][0] = dot(rX, rY);   // :95-97
    const double det = C[0][0] * C[1][1] - C[1][0] * C[0][1];
    const double id = 1.0 / det;                                      // :99 (2x2 adjugate)
    k.Ci[0][0] = C[1][1] * id; k.Ci[1][0] = -C[1][0] * id; k.Ci[0][1] = -C[0][1] * id; k.Ci[1][1] = C[0][0] * id;
    const V3 nrm = cross(rX, rY);                                     // :102-104
    const double J = norm(nrm);
    k.b = v3(nrm.x / J, nrm.y / J, nrm.z / J);
    k.E0[0][0] = 0.5 * (C[0][0] - 1.0); k.E0[1][1] = 0.5 * (C[1][1] - 1.0);   // :107
    k.E0[0][1] = k.E0[1][0] = 0.5 * C[0][1];
    k.E1[0][0] = -dot(rXX, k.b); k.E1[1][1] = -dot(rYY, k.b);         // :111-113
    k.E1[0][1] = k.E1[1][0] = -dot(rXY, k.b);
    k.rXX = rXX; k.rYY = rYY; k.rXY = rXY;
    return k;
}

// ---------------------------------------------------------------------------
// Sparsity pattern and per-GP map: plate/Assembly.h:59-153
// ---------------------------------------------------------------------------
struct Pattern {
    int nn = 0;
    std::vector<std::vector<int>> col_nodes;   // per column node: sorted rows (block level)
    std::vector<int> outer;                    // 3nn + 1 (CSC of block-upper K)
    std::vector<int> inner;                    // nnz
    std::vector<std::vector<int>> gp_map;      // per GP: one pos per local pair (i <= j)
};

static void build_outer(Pattern& p) {
    p.outer.assign(3 * p.nn + 1, 0);
    p.inner.clear();
    for (int j = 0; j < p.nn; ++j)
        for (int c = 0; c < 3; ++c) {
            for (int i : p.col_nodes[j])
                for (int r = 0; r < 3; ++r) p.inner.push_back(3 * i + r);
            p.outer[3 * j + c + 1] = static_cast<int>(p.inner.size());
        }
}

static bool init_pattern(int nn, const NbrList& nl, Pattern& p) {
    p.nn = nn;
    p.col_nodes.assign(nn, {});
    const size_t ng = nl.off.size() - 1;
    for (size_t g = 0; g < ng; ++g) {                                 // :69-75
        for (int a = nl.off[g]; a < nl.off[g + 1]; ++a)
            for (int b = nl.off[g]; b < nl.off[g + 1]; ++b) {
                const int ni = nl.ids[a], nj = nl.ids[b];
                if (ni <= nj) p.col_nodes[nj].push_back(ni);
            }
    }
    for (int j = 0; j < nn; ++j) {                                    // :78-82
        auto& c = p.col_nodes[j];
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    build_outer(p);
    p.gp_map.assign(ng, {});
    bool ok = true;
    for (size_t g = 0; g < ng; ++g) {                                 // :117-144
        const int b = nl.off[g], n = nl.off[g + 1] - b;
        auto& m = p.gp_map[g];
        m.resize(static_cast<size_t>(n) * (n + 1) / 2);
        size_t mo = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j) {
                const int ni = nl.ids[b + i], nj = nl.ids[b + j];
                const int lo = std::min(ni, nj), hi = std::max(ni, nj);
                const auto& ch = p.col_nodes[hi];
                auto it = std::lower_bound(ch.begin(), ch.end(), lo);
                if (it == ch.end() || *it != lo) { m[mo++] = -1; ok = false; }
                else m[mo++] = static_cast<int>(it - ch.begin());
            }
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Residual + tangent: plate/Assembly.h:158-345
// Per-thread F and K copies, guided(8) schedule, index-ordered merge.
// nthreads == 1 is the bit-reproducible single-thread mode.
// ---------------------------------------------------------------------------
struct Mat3x2 { V3 c0, c1; };

static void update_stiffness(const NbrList& nl, const Shape& sh, const std::vector<double>& w,
                             const ABD& mat, const Pattern& pat, const double* u,
                             std::vector<double>& F, std::vector<double>& Kv, int nthreads,
                             bool with_tangent, const std::vector<double>& kbar) {
    const int num_dofs = 3 * pat.nn;
    const int nnz = static_cast<int>(pat.inner.size());
    int T = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    T = (nthreads == 1) ? 1 : omp_get_max_threads();
#endif
    std::vector<std::vector<double>> lF(T), lK(T);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (T > 1)
#endif
    for (int t = 0; t < T; ++t) {                                     // :185-191
        lF[t].assign(num_dofs, 0.0);
        if (with_tangent) lK[t].assign(nnz, 0.0);
    }
    const int ng = static_cast<int>(nl.off.size()) - 1;
    const M3& A = mat.A; const M3& Bm_ = mat.B; const M3& D = mat.D;
#ifdef _OPENMP
#pragma omp parallel if (T > 1)
#endif
    {
        std::vector<M3> Bm, Bb, P, Q, Db;
        std::vector<double> N2x, N2y, Cix, Ciy, Tu;
#ifdef _OPENMP
#pragma omp for schedule(guided, 8)
#endif
        for (int g = 0; g < ng; ++g) {                                // :199-325
            int tid = 0;
#ifdef _OPENMP
            tid = (T > 1) ? omp_get_thread_num() : 0;
#endif
            const int b = nl.off[g], n = nl.off[g + 1] - b;
            if (n == 0) continue;
            const int* ids = nl.ids.data() + b;
            const Kin kin = kinematics(ids, n, sh.dpx.data() + b, sh.dpy.data() + b, sh.dxx.data() + b,
                                       sh.dyy.data() + b, sh.dxy.data() + b, u);
            // Constitutive::evaluate (Constitutive.h:57-70) on E1_el = E1 + kappa_bar
            const V3 e0 = v3(kin.E0[0][0], kin.E0[1][1], 2.0 * kin.E0[0][1]);
            const V3 e1 = v3(kin.E1[0][0] + kbar[g], kin.E1[1][1], 2.0 * kin.E1[0][1]);
            const V3 Nr = add(mv(A, e0), mv(Bm_, e1));
            const V3 Mr = add(mv(mT(Bm_), e0), mv(D, e1));
            const double W = w[g];
            // V = F0 * C0^{-1} (3x2)  (:217)
            Mat3x2 V;
            V.c0 = add(scl(kin.a1, kin.Ci[0][0]), scl(kin.a2, kin.Ci[1][0]));
            V.c1 = add(scl(kin.a1, kin.Ci[0][1]), scl(kin.a2, kin.Ci[1][1]));
            const double N00 = Nr.x, N01 = Nr.z, N11 = Nr.y;           // :218-219
            const V3 vc = add(add(scl(kin.rXX, Mr.x), scl(kin.rXY, 2.0 * Mr.z)), scl(kin.rYY, Mr.y));   // :221
            // p_XY = C0inv * F0^T * v_curv (:222)
            const double ft0 = dot(kin.a1, vc), ft1 = dot(kin.a2, vc);
            const double pX = kin.Ci[0][0] * ft0 + kin.Ci[0][1] * ft1;
            const double pY = kin.Ci[1][0] * ft0 + kin.Ci[1][1] * ft1;
            M3 bbT;
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) bbT.m[r][c] = comp(kin.b, r) * comp(kin.b, c);
            const double bdv = dot(kin.b, vc);                        // :224
            Bm.resize(n); Bb.resize(n); P.resize(n); Q.resize(n); Db.resize(n);
            N2x.resize(n); N2y.resize(n); Cix.resize(n); Ciy.resize(n); Tu.resize(n);
            for (int i = 0; i < n; ++i) {                             // :233-268
                const double px = sh.dpx[b + i], py = sh.dpy[b + i];
                const double hxx = sh.dxx[b + i], hyy = sh.dyy[b + i], hxy = sh.dxy[b + i];
                for (int c = 0; c < 3; ++c) {
                    Bm[i].m[0][c] = comp(kin.a1, c) * px;
                    Bm[i].m[1][c] = comp(kin.a2, c) * py;
                    Bm[i].m[2][c] = comp(kin.a1, c) * py + comp(kin.a2, c) * px;
                }
                const V3 gI = add(scl(V.c0, px), scl(V.c1, py));
                const double SXX = dot(kin.rXX, gI), SYY = dot(kin.rYY, gI), SXY = dot(kin.rXY, gI);
                for (int c = 0; c < 3; ++c) {
                    Bb[i].m[0][c] = (-hxx + SXX) * comp(kin.b, c);
                    Bb[i].m[1][c] = (-hyy + SYY) * comp(kin.b, c);
                    Bb[i].m[2][c] = 2.0 * (-hxy + SXY) * comp(kin.b, c);
                }
                const V3 fl = scl(add(mv(mT(Bm[i]), Nr), mv(mT(Bb[i]), Mr)), W);
                double* Fi = lF[tid].data() + 3 * static_cast<size_t>(ids[i]);
                Fi[0] += fl.x; Fi[1] += fl.y; Fi[2] += fl.z;
                if (!with_tangent) continue;
                P[i] = madd(mmul(A, Bm[i]), mmul(Bm_, Bb[i]));
                Q[i] = madd(mmul(Bm_, Bm[i]), mmul(D, Bb[i]));
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) Db[i].m[r][c] = -comp(gI, r) * comp(kin.b, c);
                N2x[i] = N00 * px + N01 * py; N2y[i] = N01 * px + N11 * py;
                Cix[i] = kin.Ci[0][0] * px + kin.Ci[0][1] * py;
                Ciy[i] = kin.Ci[1][0] * px + kin.Ci[1][1] * py;
                const double HM = hxx * Mr.x + 2.0 * hxy * Mr.z + hyy * Mr.y;
                const double Gp = px * pX + py * pY;
                Tu[i] = Gp - HM;
            }
            if (!with_tangent) continue;
            const std::vector<int>& map = pat.gp_map[g];
            const int* outer = pat.outer.data();
            double* Kl = lK[tid].data();
            size_t mo = 0;
            for (int i = 0; i < n; ++i) {                             // :282-324
                const M3 BmiT = mT(Bm[i]), BbiT = mT(Bb[i]), DbiT = mT(Db[i]);
                const double pxi = sh.dpx[b + i], pyi = sh.dpy[b + i];
                const double Tui = Tu[i];
                for (int j = i; j < n; ++j) {
                    const M3 Km = madd(mmul(BmiT, P[j]), mmul(BbiT, Q[j]));
                    const double geo = pxi * N2x[j] + pyi * N2y[j];
                    const M3 T12 = madd(mscl(Db[j], Tui), mscl(DbiT, Tu[j]));
                    const double t3 = bdv * (pxi * Cix[j] + pyi * Ciy[j]);
                    M3 KIJ;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c)
                            KIJ.m[r][c] = (Km.m[r][c] + (r == c ? geo : 0.0) + T12.m[r][c] + t3 * bbT.m[r][c]) * W;
                    const int ni = ids[i], nj = ids[j];
                    const int pos3 = 3 * map[mo++];
                    const bool flip = (ni > nj);
                    const int hi3 = (flip ? ni : nj) * 3;
                    for (int c = 0; c < 3; ++c) {
                        double* dst = Kl + outer[hi3 + c] + pos3;
                        if (!flip) { dst[0] += KIJ.m[0][c]; dst[1] += KIJ.m[1][c]; dst[2] += KIJ.m[2][c]; }
                        else { dst[0] += KIJ.m[c][0]; dst[1] += KIJ.m[c][1]; dst[2] += KIJ.m[c][2]; }
                    }
                }
            }
        }
    }
    F.assign(num_dofs, 0.0);                                          // :328-331
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < num_dofs; ++d) F[d] += lF[t][d];
    if (with_tangent) {
        Kv.assign(nnz, 0.0);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (T > 1)
#endif
        for (int idx = 0; idx < nnz; ++idx) {                        // :337-344
            double s = 0.0;
            for (int t = 0; t < T; ++t) s += lK[t][idx];
            Kv[idx] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Elastic hinge: plate/Constraints.h:54-119 (normal derivatives, fold angle and
// gradient) and :341-436 (evaluateElasticFold)
// ---------------------------------------------------------------------------
static void normal_derivs(V3 a1, V3 a2, V3 b, double nn, double out[3][6]) {
    double dn[3][6];
    for (int i = 0; i < 3; ++i) {
        const V3 e = v3(i == 0, i == 1, i == 2);
        const V3 c1 = cross(e, a2), c2 = cross(a1, e);
        dn[0][i] = c1.x; dn[1][i] = c1.y; dn[2][i] = c1.z;
        dn[0][i + 3] = c2.x; dn[1][i + 3] = c2.y; dn[2][i + 3] = c2.z;
    }
    double P[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P[r][c] = ((r == c ? 1.0 : 0.0) - comp(b, r) * comp(b, c)) / nn;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 6; ++c) out[r][c] = P[r][0] * dn[0][c] + P[r][1] * dn[1][c] + P[r][2] * dn[2][c];
}

static double fold_angle(const double F[12], double tx, double ty, double* grad) {
    const V3 a1L = v3(F[0], F[1], F[2]), a2L = v3(F[3], F[4], F[5]);
    const V3 a1R = v3(F[6], F[7], F[8]), a2R = v3(F[9], F[10], F[11]);
    const V3 nL = cross(a1L, a2L), nR = cross(a1R, a2R);
    const double normL = norm(nL), normR = norm(nR);
    const V3 bL = v3(nL.x / normL, nL.y / normL, nL.z / normL);
    const V3 bR = v3(nR.x / normR, nR.y / normR, nR.z / normR);
    const V3 aT = add(scl(a1L, tx), scl(a2L, ty));
    const double naT = norm(aT);
    const V3 t = v3(aT.x / naT, aT.y / naT, aT.z / naT);
    const double c = dot(bL, bR);
    const double s = dot(cross(bL, bR), t);
    const double theta = std::atan2(-s, c);
    if (grad) {
        double dbL[3][6], dbR[3][6];
        normal_derivs(a1L, a2L, bL, normL, dbL);
        normal_derivs(a1R, a2R, bR, normR, dbR);
        double daT[3][6] = {{tx, 0, 0, ty, 0, 0}, {0, tx, 0, 0, ty, 0}, {0, 0, tx, 0, 0, ty}};
        double dt[3][6];
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 6; ++cc) {
                double acc = 0;
                for (int k = 0; k < 3; ++k)
                    acc += (((r == k) ? 1.0 : 0.0) - comp(t, r) * comp(t, k)) / naT * daT[k][cc];
                dt[r][cc] = acc;
            }
        const V3 bRxt = cross(bR, t), bLxbR = cross(bL, bR), txbL = cross(t, bL);
        const double denom = c * c + s * s;
        const double idn = (denom > 1e-14) ? 1.0 / denom : 0.0;
        for (int cc = 0; cc < 6; ++cc) {
            const double dcL = bR.x * dbL[0][cc] + bR.y * dbL[1][cc] + bR.z * dbL[2][cc];
            const double dcR = bL.x * dbR[0][cc] + bL.y * dbR[1][cc] + bL.z * dbR[2][cc];
            const double dsL = (bRxt.x * dbL[0][cc] + bRxt.y * dbL[1][cc] + bRxt.z * dbL[2][cc])
                               + (bLxbR.x * dt[0][cc] + bLxbR.y * dt[1][cc] + bLxbR.z * dt[2][cc]);
            const double dsR = txbL.x * dbR[0][cc] + txbL.y * dbR[1][cc] + txbL.z * dbR[2][cc];
            grad[cc] = -(c * dsL - s * dcL) * idn;
            grad[6 + cc] = -(c * dsR - s * dcR) * idn;
        }
    }
    return theta;
}

struct Triplet { int r, c; double v; };

struct FoldDef {
    int left = -1, right = -1;      // domain indices
    double Kt = 0.0, Ang0 = 0.0;
    bool rigid = false;             // per-fold rigid hinge (constraint rows, no hinge terms)
    std::vector<double> theta0;     // per crease GP (empty: Ang0)
    std::vector<char> active;       // per crease GP (empty: all)
};

// BoundaryCondition (plate/PlateData.h:36-50); kind 1 = point domain, 2 = edge
// domain (Meshless::points / Meshless::edges), 0 = neither.
struct BC {
    int kind = 0;
    bool au = false, av = false, aw = false, arot = false;
    double vu = 0, vv = 0, vw = 0, vrot = 0;
    double force[3] = {0, 0, 0};
    double moment = 0.0;
};

struct Domain {
    std::vector<V2> X;
    std::vector<double> w;
    std::vector<V2> normal, tangent;
    NbrList nl;
    Shape sh;
    BC bc;
};

static void elastic_fold(const Domain& L, const Domain& R, const FoldDef& f, const double* u,
                         std::vector<double>& F, std::vector<Triplet>* trips, bool exact) {
    for (size_t i = 0; i < L.X.size(); ++i) {
        if (!f.active.empty() && !f.active[i]) continue;
        const int bL = L.nl.off[i], nL = L.nl.off[i + 1] - bL;
        const int bR = R.nl.off[i], nR = R.nl.off[i + 1] - bR;
        const Kin kL = kinematics(L.nl.ids.data() + bL, nL, L.sh.dpx.data() + bL, L.sh.dpy.data() + bL,
                                  L.sh.dxx.data() + bL, L.sh.dyy.data() + bL, L.sh.dxy.data() + bL, u);
        const Kin kR = kinematics(R.nl.ids.data() + bR, nR, R.sh.dpx.data() + bR, R.sh.dpy.data() + bR,
                                  R.sh.dxx.data() + bR, R.sh.dyy.data() + bR, R.sh.dxy.data() + bR, u);
        const double tx = L.tangent[i].x, ty = L.tangent[i].y;
        double Fv[12] = {kL.a1.x, kL.a1.y, kL.a1.z, kL.a2.x, kL.a2.y, kL.a2.z,
                         kR.a1.x, kR.a1.y, kR.a1.z, kR.a2.x, kR.a2.y, kR.a2.z};
        double grad[12];
        const double theta = fold_angle(Fv, tx, ty, grad);
        const double th0 = f.theta0.empty() ? f.Ang0 : f.theta0[i];
        const double mom = f.Kt * (theta - th0);
        const double W = L.w[i];
        const double Wm = W * mom, WK = W * f.Kt;
        std::vector<int> all_ids; std::vector<double> all_g; std::vector<std::array<double, 12>> Bv;
        auto extract = [&](int block, const Domain& Dm, int b, int n) {
            for (int j = 0; j < n; ++j) {
                const int id = Dm.nl.ids[b + j];
                const double dX = Dm.sh.dpx[b + j], dY = Dm.sh.dpy[b + j];
                for (int d = 0; d < 3; ++d) {
                    all_ids.push_back(id * 3 + d);
                    all_g.push_back(grad[block + d] * dX + grad[block + d + 3] * dY);
                    if (exact && trips) {
                        std::array<double, 12> bb{}; bb[block + d] = dX; bb[block + d + 3] = dY;
                        Bv.push_back(bb);
                    }
                }
            }
        };
        extract(0, L, bL, nL); extract(6, R, bR, nR);
        std::vector<std::array<double, 12>> HB;
        if (exact && trips && std::fabs(Wm) > 0.0) {                 // :402-416
            double H[12][12];
            double fmax = 0; for (int k = 0; k < 12; ++k) fmax = std::max(fmax, std::fabs(Fv[k]));
            const double hs = 1e-6 * std::max(1.0, fmax);
            for (int k = 0; k < 12; ++k) {
                double Fp[12], Fm[12], gp[12], gm[12];
                std::memcpy(Fp, Fv, sizeof(Fv)); std::memcpy(Fm, Fv, sizeof(Fv));
                Fp[k] += hs; Fm[k] -= hs;
                fold_angle(Fp, tx, ty, gp); fold_angle(Fm, tx, ty, gm);
                for (int r = 0; r < 12; ++r) H[r][k] = (gp[r] - gm[r]) / (2.0 * hs);
            }
            double Hs[12][12];
            for (int r = 0; r < 12; ++r) for (int c = 0; c < 12; ++c) Hs[r][c] = 0.5 * (H[r][c] + H[c][r]);
            HB.resize(Bv.size());
            for (size_t m = 0; m < Bv.size(); ++m)
                for (int r = 0; r < 12; ++r) {
                    double acc = 0; for (int c = 0; c < 12; ++c) acc += Hs[r][c] * Bv[m][c];
                    HB[m][r] = acc;
                }
        }
        const bool use_h = !HB.empty();
        for (size_t k = 0; k < all_ids.size(); ++k) {                 // :420-434
            F[all_ids[k]] += all_g[k] * Wm;
            if (!trips) continue;
            for (size_t m = 0; m < all_ids.size(); ++m) {
                double val = all_g[k] * all_g[m] * WK;
                if (use_h) { double d = 0; for (int r = 0; r < 12; ++r) d += Bv[k][r] * HB[m][r]; val += Wm * d; }
                if (all_ids[k] / 3 <= all_ids[m] / 3) trips->push_back({all_ids[k], all_ids[m], val});
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Constraint rows: Constraints::evaluateDirichlet (plate/Constraints.h:24-52)
// and the pointwise tie / rigid angle rows of Constraints::evaluateFold
// (plate/Constraints.h:247-308). Triplets are pushed in the reference order.
// ---------------------------------------------------------------------------
static void dirichlet_rows(const Domain& d, const double* u, int& row, std::vector<double>& c,
                           std::vector<Triplet>& J) {
    const BC& bc = d.bc;
    for (size_t g = 0; g < d.X.size(); ++g) {
        const int b = d.nl.off[g], n = d.nl.off[g + 1] - b;
        double uc[3] = {0, 0, 0}, ux[3] = {0, 0, 0}, uy[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {                                 // :31-36
            const double* uj = u + 3 * static_cast<size_t>(d.nl.ids[b + j]);
            for (int k = 0; k < 3; ++k) {
                uc[k] += uj[k] * d.sh.psi[b + j];
                ux[k] += uj[k] * d.sh.dpx[b + j];
                uy[k] += uj[k] * d.sh.dpy[b + j];
            }
        }
        const bool act[3] = {bc.au, bc.av, bc.aw};
        const double val[3] = {bc.vu, bc.vv, bc.vw};
        for (int k = 0; k < 3; ++k) {                                 // :38-40
            if (!act[k]) continue;
            c.push_back(uc[k] - val[k]);
            for (int j = 0; j < n; ++j) J.push_back({row, d.nl.ids[b + j] * 3 + k, d.sh.psi[b + j]});
            ++row;
        }
        if (bc.arot) {                                                // :42-48
            const double nx = d.normal[g].x, ny = d.normal[g].y;
            c.push_back((ux[2] * nx + uy[2] * ny) - bc.vrot);
            for (int j = 0; j < n; ++j)
                J.push_back({row, d.nl.ids[b + j] * 3 + 2, d.sh.dpx[b + j] * nx + d.sh.dpy[b + j] * ny});
            ++row;
        }
    }
}

static void fold_rows(const Domain& L, const Domain& R, const FoldDef& f, const double* u, bool rig,
                      int& row, std::vector<double>& c, std::vector<Triplet>& J) {
    for (size_t i = 0; i < L.X.size(); ++i) {
        if (!f.active.empty() && !f.active[i]) continue;
        const int bL = L.nl.off[i], nL = L.nl.off[i + 1] - bL;
        const int bR = R.nl.off[i], nR = R.nl.off[i + 1] - bR;
        double uL[3] = {0, 0, 0}, uR[3] = {0, 0, 0};                  // :267-269
        for (int j = 0; j < nL; ++j)
            for (int k = 0; k < 3; ++k) uL[k] += u[3 * static_cast<size_t>(L.nl.ids[bL + j]) + k] * L.sh.psi[bL + j];
        for (int j = 0; j < nR; ++j)
            for (int k = 0; k < 3; ++k) uR[k] += u[3 * static_cast<size_t>(R.nl.ids[bR + j]) + k] * R.sh.psi[bR + j];
        for (int d = 0; d < 3; ++d) {                                 // :271-278
            c.push_back(uL[d] - uR[d]);
            for (int j = 0; j < nL; ++j) J.push_back({row, L.nl.ids[bL + j] * 3 + d, L.sh.psi[bL + j]});
            for (int j = 0; j < nR; ++j) J.push_back({row, R.nl.ids[bR + j] * 3 + d, -R.sh.psi[bR + j]});
            ++row;
        }
        if (!rig) continue;
        const Kin kL = kinematics(L.nl.ids.data() + bL, nL, L.sh.dpx.data() + bL, L.sh.dpy.data() + bL,
                                  L.sh.dxx.data() + bL, L.sh.dyy.data() + bL, L.sh.dxy.data() + bL, u);
        const Kin kR = kinematics(R.nl.ids.data() + bR, nR, R.sh.dpx.data() + bR, R.sh.dpy.data() + bR,
                                  R.sh.dxx.data() + bR, R.sh.dyy.data() + bR, R.sh.dxy.data() + bR, u);
        double Fv[12] = {kL.a1.x, kL.a1.y, kL.a1.z, kL.a2.x, kL.a2.y, kL.a2.z,
                         kR.a1.x, kR.a1.y, kR.a1.z, kR.a2.x, kR.a2.y, kR.a2.z};
        double grad[12];
        const double theta = fold_angle(Fv, L.tangent[i].x, L.tangent[i].y, grad);   // :280-287
        c.push_back(theta - (f.theta0.empty() ? f.Ang0 : f.theta0[i]));
        auto add = [&](int block, const Domain& Dm, int b, int n) {    // :293-301
            for (int j = 0; j < n; ++j) {
                const double dX = Dm.sh.dpx[b + j], dY = Dm.sh.dpy[b + j];
                for (int d = 0; d < 3; ++d)
                    J.push_back({row, Dm.nl.ids[b + j] * 3 + d, grad[block + d] * dX + grad[block + d + 3] * dY});
            }
        };
        add(0, L, bL, nL); add(6, R, bR, nR);
        ++row;
    }
}

// ---------------------------------------------------------------------------
// The plate: plate/Meshless.h:11-127 (domains, folds, initAssembly,
// assembleSystem, assembleSystemInternalOnly)
// ---------------------------------------------------------------------------
struct Plate {
    std::vector<V2> nodes;
    std::vector<Domain> doms;       // dom 0 is the face domain by convention of the caller
    int face = -1;
    std::vector<FoldDef> folds;
    std::vector<FoldLine> creases;
    ABD mat;
    NatCurv nat;
    double rho = 0.0;
    int order = 2, kernel = 0;
    Pattern pat;
    // last assembled system (CSC of block-upper K)
    std::vector<int> K_outer, K_inner;
    std::vector<double> K_val, F;
    // last evaluated constraints
    std::vector<double> c;
    std::vector<Triplet> J;
    std::string err;
    double t_last = 0.0;
};

}  // namespace orc

using namespace orc;

// ============================================================================
// C interface for ctypes (tests and the CPU baseline)
// ============================================================================
extern "C" {

Plate* orc_create() { return new Plate(); }
void orc_destroy(Plate* p) { delete p; }

// Kinematics::naturalCurvature() (plate/Kinematics.h:42-45), set from the
// configuration's "natural_curvature" band (main.cpp:80-97).
void orc_set_natural_curvature(Plate* p, int active, double x0, double width, double theta0, double sign) {
    p->nat.active = active != 0;
    p->nat.x0 = x0; p->nat.width = width; p->nat.theta0 = theta0; p->nat.sign = sign;
}
const char* orc_last_error(Plate* p) { return p->err.c_str(); }

void orc_set_nodes(Plate* p, int nn, const double* X, const double* Y) {
    p->nodes.resize(nn);
    for (int i = 0; i < nn; ++i) { p->nodes[i].x = X[i]; p->nodes[i].y = Y[i]; }
}

void orc_set_material_plies(Plate* p, int nply, const double* plies /* nply x 6 */) {
    std::vector<Ply> v(nply);
    for (int i = 0; i < nply; ++i) {
        v[i].E1 = plies[6 * i + 0]; v[i].E2 = plies[6 * i + 1]; v[i].nu12 = plies[6 * i + 2];
        v[i].G12 = plies[6 * i + 3]; v[i].angle_deg = plies[6 * i + 4]; v[i].thickness = plies[6 * i + 5];
    }
    p->mat = computeABD(v);
}

void orc_get_abd(Plate* p, double* A, double* B, double* D) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            A[3 * r + c] = p->mat.A.m[r][c]; B[3 * r + c] = p->mat.B.m[r][c]; D[3 * r + c] = p->mat.D.m[r][c];
        }
}

void orc_add_crease(Plate* p, int nv, const double* vx, const double* vy, int ext_s, int ext_e, int closed) {
    FoldLine f;
    f.v.resize(nv);
    for (int i = 0; i < nv; ++i) { f.v[i].x = vx[i]; f.v[i].y = vy[i]; }
    f.ext_s = ext_s != 0; f.ext_e = ext_e != 0; f.closed = closed != 0;
    p->creases.push_back(f);
}

int orc_add_domain(Plate* p, int ng, const double* gx, const double* gy, const double* w,
                   const double* nx, const double* ny, const double* tx, const double* ty, int is_face) {
    Domain d;
    d.X.resize(ng); d.w.resize(ng); d.normal.resize(ng); d.tangent.resize(ng);
    for (int i = 0; i < ng; ++i) {
        d.X[i].x = gx[i]; d.X[i].y = gy[i]; d.w[i] = w[i];
        d.normal[i].x = nx ? nx[i] : 0.0; d.normal[i].y = ny ? ny[i] : 0.0;
        d.tangent[i].x = tx ? tx[i] : 0.0; d.tangent[i].y = ty ? ty[i] : 0.0;
    }
    p->doms.push_back(std::move(d));
    const int id = static_cast<int>(p->doms.size()) - 1;
    if (is_face) p->face = id;
    return id;
}

int orc_add_fold(Plate* p, int left, int right, double Kt, double Ang0) {
    FoldDef f; f.left = left; f.right = right; f.Kt = Kt; f.Ang0 = Ang0;
    p->folds.push_back(f);
    return static_cast<int>(p->folds.size()) - 1;
}

void orc_set_fold_rigid(Plate* p, int fold, int rigid) { p->folds[fold].rigid = rigid != 0; }

void orc_set_fold_gp_state(Plate* p, int fold, int n, const double* theta0, const char* active) {
    FoldDef& f = p->folds[fold];
    f.theta0.assign(theta0, theta0 + (theta0 ? n : 0));
    f.active.assign(active, active + (active ? n : 0));
}

// Indicator of one point against crease c: out = f, d, gx, gy, hxx, hyy, hxy, valid
void orc_indicator(Plate* p, int c, double x, double y, double* out) {
    V2 P; P.x = x; P.y = y;
    const Ind r = indicator(P, p->creases[c]);
    out[0] = r.f; out[1] = r.d; out[2] = r.gx; out[3] = r.gy; out[4] = r.hxx; out[5] = r.hyy; out[6] = r.hxy;
    out[7] = r.valid ? 1.0 : 0.0;
}

void orc_window(int kernel, int n, const double* s, double* w, double* dw, double* d2w) {
    for (int i = 0; i < n; ++i) { w[i] = kw(kernel, s[i]); dw[i] = kdw(kernel, s[i]); d2w[i] = kd2w(kernel, s[i]); }
}

// FRNN for one domain (FRNN.h:23-78). Returns the number of pairs.
long orc_neighbours(Plate* p, int dom, double rho) {
    p->rho = rho;
    Domain& d = p->doms[dom];
    frnn(p->nodes, d.X, rho, d.nl);
    return static_cast<long>(d.nl.ids.size());
}

void orc_get_neighbours(Plate* p, int dom, int* off, int* ids, double* s) {
    const Domain& d = p->doms[dom];
    std::memcpy(off, d.nl.off.data(), d.nl.off.size() * sizeof(int));
    std::memcpy(ids, d.nl.ids.data(), d.nl.ids.size() * sizeof(int));
    std::memcpy(s, d.nl.s.data(), d.nl.s.size() * sizeof(double));
}

// Replace a domain's neighbour list (used to feed identical lists to both sides).
void orc_set_neighbours(Plate* p, int dom, int ng, const int* off, const int* ids, const double* s) {
    Domain& d = p->doms[dom];
    d.nl.off.assign(off, off + ng + 1);
    d.nl.ids.assign(ids, ids + off[ng]);
    d.nl.s.assign(s, s + off[ng]);
}

// GlobalRKPM::solve for one domain. Returns 0, 3 (singular) or 4 (pivot).
int orc_shape_functions(Plate* p, int dom, double rho, int order, int kernel, int nthreads, int* nbad, int* first) {
    Domain& d = p->doms[dom];
    p->rho = rho; p->order = order; p->kernel = kernel;
    const RkpmStatus st = rkpm(p->nodes, d.X, d.nl, rho, order, kernel, p->creases, d.sh, nthreads);
    if (nbad) *nbad = st.nbad;
    if (first) *first = st.first;
    return st.code;
}

void orc_get_shape_functions(Plate* p, int dom, double* psi, double* dpx, double* dpy, double* dxx, double* dyy, double* dxy) {
    const Domain& d = p->doms[dom];
    const size_t n = d.sh.psi.size() * sizeof(double);
    std::memcpy(psi, d.sh.psi.data(), n); std::memcpy(dpx, d.sh.dpx.data(), n); std::memcpy(dpy, d.sh.dpy.data(), n);
    std::memcpy(dxx, d.sh.dxx.data(), n); std::memcpy(dyy, d.sh.dyy.data(), n); std::memcpy(dxy, d.sh.dxy.data(), n);
}

// NonLinearAssembler::initSparsityPattern over the face domain. Returns 0 or 5 (map miss).
int orc_init_assembly(Plate* p) {
    const bool ok = init_pattern(static_cast<int>(p->nodes.size()), p->doms[p->face].nl, p->pat);
    return ok ? 0 : 5;
}

// Face pattern (before any fold union): sizes
long orc_face_nnz(Plate* p) { return static_cast<long>(p->pat.inner.size()); }
void orc_get_face_pattern(Plate* p, int* outer, int* inner) {
    std::memcpy(outer, p->pat.outer.data(), p->pat.outer.size() * sizeof(int));
    std::memcpy(inner, p->pat.inner.data(), p->pat.inner.size() * sizeof(int));
}
void orc_get_gp_map(Plate* p, int g, int* out) {
    const auto& m = p->pat.gp_map[g];
    std::memcpy(out, m.data(), m.size() * sizeof(int));
}

// Meshless::assembleSystem (Meshless.h:113-127). Stores K (union pattern) and
// F_int; returns nnz of K. nthreads = 1 is the single-thread oracle mode.
long orc_assemble(Plate* p, const double* x, int flag_rig, int exact_tangent, int nthreads) {
    Domain& fd = p->doms[p->face];
    std::vector<double> Kv;
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
    update_stiffness(fd.nl, fd.sh, fd.w, p->mat, p->pat, x, p->F, Kv, nthreads, true, natural_kappa(p->nat, fd.X));
    const int nn = p->pat.nn;
    if (flag_rig || p->folds.empty()) {
        p->K_outer = p->pat.outer; p->K_inner = p->pat.inner; p->K_val = Kv;
    } else {
        std::vector<Triplet> trips;
        for (const auto& f : p->folds)
            if (!f.rigid) elastic_fold(p->doms[f.left], p->doms[f.right], f, x, p->F, &trips, exact_tangent != 0);
        // setFromTriplets (duplicates summed in triplet order), then K + K_fold
        // over the union of both patterns.
        std::vector<std::map<int, double>> cols(3 * nn);
        for (const auto& t : trips) cols[t.c][t.r] += t.v;
        p->K_outer.assign(3 * nn + 1, 0);
        p->K_inner.clear(); p->K_val.clear();
        for (int c = 0; c < 3 * nn; ++c) {
            auto it = cols[c].begin();
            int a = p->pat.outer[c];
            const int ae = p->pat.outer[c + 1];
            while (a < ae || it != cols[c].end()) {
                if (it == cols[c].end() || (a < ae && p->pat.inner[a] < it->first)) {
                    p->K_inner.push_back(p->pat.inner[a]); p->K_val.push_back(Kv[a]); ++a;
                } else if (a >= ae || it->first < p->pat.inner[a]) {
                    p->K_inner.push_back(it->first); p->K_val.push_back(it->second); ++it;
                } else {
                    p->K_inner.push_back(p->pat.inner[a]); p->K_val.push_back(Kv[a] + it->second); ++a; ++it;
                }
            }
            p->K_outer[c + 1] = static_cast<int>(p->K_inner.size());
        }
    }
#ifdef _OPENMP
    p->t_last = omp_get_wtime() - t0;
#endif
    return static_cast<long>(p->K_inner.size());
}

// Meshless::assembleSystemInternalOnly (Meshless.h:102-111): the reference runs
// the full updateStiffness here too; with_tangent = 1 reproduces that cost.
void orc_internal_force(Plate* p, const double* x, int flag_rig, int with_tangent, int nthreads, double* F_out) {
    Domain& fd = p->doms[p->face];
    std::vector<double> Kv, F;
    update_stiffness(fd.nl, fd.sh, fd.w, p->mat, p->pat, x, F, Kv, nthreads, with_tangent != 0,
                     natural_kappa(p->nat, fd.X));
    if (!flag_rig)
        for (const auto& f : p->folds)
            if (!f.rigid) elastic_fold(p->doms[f.left], p->doms[f.right], f, x, F, nullptr, false);
    std::memcpy(F_out, F.data(), F.size() * sizeof(double));
}

void orc_get_system(Plate* p, int* outer, int* inner, double* val, double* F) {
    std::memcpy(outer, p->K_outer.data(), p->K_outer.size() * sizeof(int));
    std::memcpy(inner, p->K_inner.data(), p->K_inner.size() * sizeof(int));
    std::memcpy(val, p->K_val.data(), p->K_val.size() * sizeof(double));
    std::memcpy(F, p->F.data(), p->F.size() * sizeof(double));
}

double orc_last_seconds(Plate* p) { return p->t_last; }

// Fold angle and its 12-gradient (Constraints.h:71-119), for closed-form checks.
double orc_fold_angle(const double* F, double tx, double ty, double* grad) {
    return fold_angle(F, tx, ty, grad);
}

// Per-GP kinematics of a domain (Kinematics.h:61-120): out per GP =
// a1(3) a2(3) b(3) E0(xx,yy,xy) E1(xx,yy,xy) = 15 doubles.
void orc_kinematics(Plate* p, int dom, const double* x, double* out) {
    const Domain& d = p->doms[dom];
    for (size_t g = 0; g + 1 < d.nl.off.size(); ++g) {
        const int b = d.nl.off[g], n = d.nl.off[g + 1] - b;
        const Kin k = kinematics(d.nl.ids.data() + b, n, d.sh.dpx.data() + b, d.sh.dpy.data() + b,
                                 d.sh.dxx.data() + b, d.sh.dyy.data() + b, d.sh.dxy.data() + b, x);
        double* o = out + 15 * g;
        o[0] = k.a1.x; o[1] = k.a1.y; o[2] = k.a1.z; o[3] = k.a2.x; o[4] = k.a2.y; o[5] = k.a2.z;
        o[6] = k.b.x; o[7] = k.b.y; o[8] = k.b.z;
        o[9] = k.E0[0][0]; o[10] = k.E0[1][1]; o[11] = k.E0[0][1];
        o[12] = k.E1[0][0]; o[13] = k.E1[1][1]; o[14] = k.E1[0][1];
    }
}

// Meshless::computeStrainEnergies (plate/Meshless.h:205-219) over the face
// domain and computeTotalHingeEnergy (:198-203, Constraints::computeHingeEnergy,
// plate/Constraints.h:310-339) over every fold's active crease GPs, summed in
// GP order. out = E_mem, E_ben, E_hin.
void orc_energies(Plate* p, const double* x, double* out) {
    const Domain& d = p->doms[p->face];
    const ABD& m = p->mat;
    const std::vector<double> kbar = natural_kappa(p->nat, d.X);
    double Em = 0.0, Eb = 0.0;
    for (size_t g = 0; g + 1 < d.nl.off.size(); ++g) {
        const int b = d.nl.off[g], n = d.nl.off[g + 1] - b;
        const Kin k = kinematics(d.nl.ids.data() + b, n, d.sh.dpx.data() + b, d.sh.dpy.data() + b,
                                 d.sh.dxx.data() + b, d.sh.dyy.data() + b, d.sh.dxy.data() + b, x);
        const V3 e0 = v3(k.E0[0][0], k.E0[1][1], 2.0 * k.E0[0][1]);
        const V3 e1 = v3(k.E1[0][0] + kbar[g], k.E1[1][1], 2.0 * k.E1[0][1]);   // E1_el (Meshless.h:212)
        const V3 N = add(mv(m.A, e0), mv(m.B, e1));                       // Constitutive.h:67-68
        const V3 M = add(mv(mT(m.B), e0), mv(m.D, e1));
        Em += 0.5 * d.w[g] * dot(e0, N);
        Eb += 0.5 * d.w[g] * dot(e1, M);
    }
    double Eh = 0.0;
    for (const FoldDef& f : p->folds) {
        const Domain& L = p->doms[f.left];
        const Domain& R = p->doms[f.right];
        for (size_t i = 0; i < L.X.size(); ++i) {
            if (!f.active.empty() && !f.active[i]) continue;
            double F[12];
            for (int side = 0; side < 2; ++side) {
                const Domain& D = side ? R : L;
                const int b = D.nl.off[i], n = D.nl.off[i + 1] - b;
                const Kin k = kinematics(D.nl.ids.data() + b, n, D.sh.dpx.data() + b, D.sh.dpy.data() + b,
                                         D.sh.dxx.data() + b, D.sh.dyy.data() + b, D.sh.dxy.data() + b, x);
                F[6 * side + 0] = k.a1.x; F[6 * side + 1] = k.a1.y; F[6 * side + 2] = k.a1.z;
                F[6 * side + 3] = k.a2.x; F[6 * side + 4] = k.a2.y; F[6 * side + 5] = k.a2.z;
            }
            const double th = fold_angle(F, L.tangent[i].x, L.tangent[i].y, nullptr);
            const double dth = th - (f.theta0.empty() ? f.Ang0 : f.theta0[i]);
            Eh += L.w[i] * (0.5 * f.Kt * (dth * dth));
        }
    }
    out[0] = Em; out[1] = Eb; out[2] = Eh;
}

// Dihedral angle at every crease GP of a fold (Constraints::foldAngleAndGrad,
// plate/Constraints.h:71-88), active or not.
void orc_fold_angles(Plate* p, int fold, const double* x, double* theta) {
    const FoldDef& f = p->folds[fold];
    const Domain& L = p->doms[f.left];
    const Domain& R = p->doms[f.right];
    for (size_t i = 0; i < L.X.size(); ++i) {
        double F[12];
        for (int side = 0; side < 2; ++side) {
            const Domain& D = side ? R : L;
            const int b = D.nl.off[i], n = D.nl.off[i + 1] - b;
            const Kin k = kinematics(D.nl.ids.data() + b, n, D.sh.dpx.data() + b, D.sh.dpy.data() + b,
                                     D.sh.dxx.data() + b, D.sh.dyy.data() + b, D.sh.dxy.data() + b, x);
            F[6 * side + 0] = k.a1.x; F[6 * side + 1] = k.a1.y; F[6 * side + 2] = k.a1.z;
            F[6 * side + 3] = k.a2.x; F[6 * side + 4] = k.a2.y; F[6 * side + 5] = k.a2.z;
        }
        theta[i] = fold_angle(F, L.tangent[i].x, L.tangent[i].y, nullptr);
    }
}

// Removes every crease; shape functions must be recomputed.
void orc_clear_creases(Plate* p) { p->creases.clear(); }

void orc_set_domain_bc(Plate* p, int dom, int kind, int active, const double* val, const double* force,
                       double moment) {
    BC& bc = p->doms[dom].bc;
    bc.kind = kind;
    bc.au = (active & 1) != 0; bc.av = (active & 2) != 0; bc.aw = (active & 4) != 0; bc.arot = (active & 8) != 0;
    bc.vu = val[0]; bc.vv = val[1]; bc.vw = val[2]; bc.vrot = val[3];
    for (int k = 0; k < 3; ++k) bc.force[k] = force[k];
    bc.moment = moment;
}

// Meshless::evaluateConstraints (plate/Meshless.h:129-195): Dirichlet rows of the
// point domains, then of the edge domains, then each fold's rows. Returns the
// triplet count; the row count goes to *nrows.
long orc_constraints(Plate* p, const double* x, int flag_rig, long* nrows) {
    p->c.clear(); p->J.clear();
    int row = 0;
    for (int kind = 1; kind <= 2; ++kind)
        for (const Domain& d : p->doms)
            if (d.bc.kind == kind && (d.bc.au || d.bc.av || d.bc.aw || d.bc.arot))
                dirichlet_rows(d, x, row, p->c, p->J);
    for (const auto& f : p->folds)
        fold_rows(p->doms[f.left], p->doms[f.right], f, x, flag_rig != 0 || f.rigid, row, p->c, p->J);
    *nrows = row;
    return static_cast<long>(p->J.size());
}

void orc_get_constraints(Plate* p, double* c, int* rows, int* cols, double* vals) {
    std::memcpy(c, p->c.data(), p->c.size() * sizeof(double));
    for (size_t k = 0; k < p->J.size(); ++k) { rows[k] = p->J[k].r; cols[k] = p->J[k].c; vals[k] = p->J[k].v; }
}

// Meshless::computeExternalForce (plate/Meshless.h:60-100): dead forces of the
// point, edge and fold domains and the edge moments, accumulated in the
// reference's order (points, then per edge force then moment, then folds).
void orc_external_force(Plate* p, double* F) {
    const size_t nd = 3 * p->nodes.size();
    for (size_t k = 0; k < nd; ++k) F[k] = 0.0;
    auto force = [&](const Domain& d) {
        const double* fv = d.bc.force;
        if (fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2] <= 0.0) return;
        for (size_t g = 0; g < d.X.size(); ++g) {
            const double W = d.w[g];
            for (int j = d.nl.off[g]; j < d.nl.off[g + 1]; ++j) {
                const double pw = d.sh.psi[j] * W;
                for (int k = 0; k < 3; ++k) F[3 * static_cast<size_t>(d.nl.ids[j]) + k] += pw * fv[k];
            }
        }
    };
    auto moment = [&](const Domain& d) {
        if (d.bc.moment == 0.0) return;
        for (size_t g = 0; g < d.X.size(); ++g) {
            double side = 0.0;                                        // :86-88
            for (int j = d.nl.off[g]; j < d.nl.off[g + 1]; ++j) {
                const V2 X = p->nodes[d.nl.ids[j]];
                side += (X.x - d.X[g].x) * d.normal[g].x + (X.y - d.X[g].y) * d.normal[g].y;
            }
            const double nx = side > 0.0 ? -d.normal[g].x : d.normal[g].x;
            const double ny = side > 0.0 ? -d.normal[g].y : d.normal[g].y;
            for (int j = d.nl.off[g]; j < d.nl.off[g + 1]; ++j)
                F[3 * static_cast<size_t>(d.nl.ids[j]) + 2] +=
                    d.w[g] * d.bc.moment * (d.sh.dpx[j] * nx + d.sh.dpy[j] * ny);
        }
    };
    for (const Domain& d : p->doms) if (d.bc.kind == 1) force(d);
    for (const Domain& d : p->doms) if (d.bc.kind == 2) { force(d); moment(d); }
    for (const auto& f : p->folds) { force(p->doms[f.left]); force(p->doms[f.right]); }
}

}  // extern "C"
