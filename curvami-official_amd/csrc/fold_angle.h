// Fold angle between the two one-sided tangent planes at a crease Gauss point
// and its gradient with respect to the 12 deformation-gradient entries
// F = [a1_L, a2_L, a1_R, a2_R]: Constraints::computeNormalDerivatives and
// foldAngleAndGrad (plate/Constraints.h:54-119). theta = atan2(-(b_L x b_R) . t,
// b_L . b_R), t the unit image of the crease tangent on the left side.
#pragma once
#include <hip/hip_runtime.h>

namespace cvm {

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline void normal_derivs(const double* a1, const double* a2, const double* b, double nn, double out[3][6]) {
    double dn[3][6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double e[3] = {0, 0, 0};
        e[i] = 1.0;
        double c1[3], c2[3];
        cross3(e, a2, c1);
        cross3(a1, e, c2);
        dn[0][i] = c1[0]; dn[1][i] = c1[1]; dn[2][i] = c1[2];
        dn[0][i + 3] = c2[0]; dn[1][i + 3] = c2[1]; dn[2][i + 3] = c2[2];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += (((r == k) ? 1.0 : 0.0) - b[r] * b[k]) / nn * dn[k][c];
            out[r][c] = acc;
        }
}

// Unit normals, unit crease tangent (left side) and the angle's cosine / sine.
struct FoldFrame {
    double normL, normR, naT;
    double bL[3], bR[3], t[3], bLxbR[3];
    double c, s;
};

__device__ inline double fold_frame(const double* F, double tx, double ty, FoldFrame& f) {
    const double* a1L = F; const double* a2L = F + 3; const double* a1R = F + 6; const double* a2R = F + 9;
    double nL[3], nR[3];
    cross3(a1L, a2L, nL);
    cross3(a1R, a2R, nR);
    f.normL = sqrt(nL[0] * nL[0] + nL[1] * nL[1] + nL[2] * nL[2]);
    f.normR = sqrt(nR[0] * nR[0] + nR[1] * nR[1] + nR[2] * nR[2]);
    for (int k = 0; k < 3; ++k) { f.bL[k] = nL[k] / f.normL; f.bR[k] = nR[k] / f.normR; }
    const double aT[3] = {a1L[0] * tx + a2L[0] * ty, a1L[1] * tx + a2L[1] * ty, a1L[2] * tx + a2L[2] * ty};
    f.naT = sqrt(aT[0] * aT[0] + aT[1] * aT[1] + aT[2] * aT[2]);
    for (int k = 0; k < 3; ++k) f.t[k] = aT[k] / f.naT;
    f.c = f.bL[0] * f.bR[0] + f.bL[1] * f.bR[1] + f.bL[2] * f.bR[2];
    cross3(f.bL, f.bR, f.bLxbR);
    f.s = f.bLxbR[0] * f.t[0] + f.bLxbR[1] * f.t[1] + f.bLxbR[2] * f.t[2];
    return atan2(-f.s, f.c);
}

__device__ inline double fold_angle(const double* F, double tx, double ty) {
    FoldFrame f;
    return fold_frame(F, tx, ty, f);
}

__device__ inline double fold_angle_grad(const double* F, double tx, double ty, double* grad) {
    const double* a1L = F; const double* a2L = F + 3; const double* a1R = F + 6; const double* a2R = F + 9;
    FoldFrame f;
    const double theta = fold_frame(F, tx, ty, f);
    const double normL = f.normL, normR = f.normR, naT = f.naT, c = f.c, s = f.s;
    const double* bL = f.bL; const double* bR = f.bR; const double* t = f.t; const double* bLxbR = f.bLxbR;
    double dbL[3][6], dbR[3][6];
    normal_derivs(a1L, a2L, bL, normL, dbL);
    normal_derivs(a1R, a2R, bR, normR, dbR);
    const double daT[3][6] = {{tx, 0, 0, ty, 0, 0}, {0, tx, 0, 0, ty, 0}, {0, 0, tx, 0, 0, ty}};
    double dt[3][6];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += (((r == k) ? 1.0 : 0.0) - t[r] * t[k]) / naT * daT[k][cc];
            dt[r][cc] = acc;
        }
    double bRxt[3], txbL[3];
    cross3(bR, t, bRxt);
    cross3(t, bL, txbL);
    const double denom = c * c + s * s;
    const double idn = (denom > 1e-14) ? 1.0 / denom : 0.0;
#pragma unroll
    for (int cc = 0; cc < 6; ++cc) {
        const double dcL = bR[0] * dbL[0][cc] + bR[1] * dbL[1][cc] + bR[2] * dbL[2][cc];
        const double dcR = bL[0] * dbR[0][cc] + bL[1] * dbR[1][cc] + bL[2] * dbR[2][cc];
        const double dsL = (bRxt[0] * dbL[0][cc] + bRxt[1] * dbL[1][cc] + bRxt[2] * dbL[2][cc])
                         + (bLxbR[0] * dt[0][cc] + bLxbR[1] * dt[1][cc] + bLxbR[2] * dt[2][cc]);
        const double dsR = txbL[0] * dbR[0][cc] + txbL[1] * dbR[1][cc] + txbL[2] * dbR[2][cc];
        grad[cc] = -(c * dsL - s * dcL) * idn;
        grad[6 + cc] = -(c * dsR - s * dcR) * idn;
    }
    return theta;
}

}  // namespace cvm
