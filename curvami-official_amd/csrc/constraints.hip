// Constraint rows and the external reference load, on the device, FP64.
//
//  * Constraint rows (Meshless::evaluateConstraints, plate/Meshless.h:129-195):
//    Dirichlet rows of the point domains, then of the edge domains
//    (Constraints::evaluateDirichlet, plate/Constraints.h:24-52), then per fold
//    and active crease GP the pointwise tie rows u_L - u_R (3 rows) and, with
//    rigid folds, the angle row theta - theta0 (Constraints::evaluateFold,
//    plate/Constraints.h:247-308). The triplet order is the reference's push
//    order, so J built from (rows, cols, values) sums the same duplicates.
//    One thread per job (a Dirichlet GP or a crease GP): these rows touch only
//    boundary and crease GPs, a few thousand at most.
//  * External load (Meshless::computeExternalForce, plate/Meshless.h:60-100):
//    dead forces (Psi W) f on point, edge and fold domains and edge moments
//    W m (grad Psi . n_out). The host lists every contribution of a dof in the
//    reference's accumulation order; one thread per dof sums its list.
//
// Rounding: both kernels evaluate with contraction off, term by term in the
// reference's summation order; the remaining differences from a CPU evaluation
// come from the shape functions themselves (round-off level).
#include <algorithm>

#include "cvm_internal.h"
#include "fold_angle.h"

namespace cvm {

namespace {

struct CDom {
    const double* sf;        // 6 planes of npairs
    const int* off;
    const int* ids;
    int64_t np;
    const double* nx;
    const double* ny;
    const double* tx;
    const double* ty;
    const double* w;
    double val[4];
    int active;
};

struct CFold {
    int left, right;
    const double* theta0;    // may be null
    double Ang0;
    int rigid;               // angle row for this fold
};

enum : int { JOB_DIRICHLET = 0, JOB_FOLD = 1 };
constexpr int JOB_INTS = 6;

}  // namespace

#pragma clang fp contract(off)

__global__ void k_constraints(int njobs, const int* __restrict__ jobs, const CDom* __restrict__ doms,
                              const CFold* __restrict__ folds, int rigid, const double* __restrict__ u,
                              double* __restrict__ cvals, double* __restrict__ jvals) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= njobs) return;
    const int* jb = jobs + JOB_INTS * q;
    const int kind = jb[0], gp = jb[3], row0 = jb[4];
    int64_t t = jb[5];
    if (kind == JOB_DIRICHLET) {
        const CDom d = doms[jb[1]];
        const int b = d.off[gp], e = d.off[gp + 1];
        const double* psi = d.sf;
        const double* dpx = d.sf + d.np;
        const double* dpy = d.sf + 2 * d.np;
        double uc[3] = {0.0, 0.0, 0.0}, ux[3] = {0.0, 0.0, 0.0}, uy[3] = {0.0, 0.0, 0.0};
        for (int k = b; k < e; ++k) {
            const double* uk = u + 3 * static_cast<int64_t>(d.ids[k]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                uc[c] = uc[c] + uk[c] * psi[k];
                ux[c] = ux[c] + uk[c] * dpx[k];
                uy[c] = uy[c] + uk[c] * dpy[k];
            }
        }
        int row = row0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!(d.active & (1 << c))) continue;
            cvals[row++] = uc[c] - d.val[c];
            for (int k = b; k < e; ++k) jvals[t++] = psi[k];
        }
        if (d.active & 8) {
            const double nx = d.nx[gp], ny = d.ny[gp];
            cvals[row] = (ux[2] * nx + uy[2] * ny) - d.val[3];
            for (int k = b; k < e; ++k) jvals[t++] = dpx[k] * nx + dpy[k] * ny;
        }
        return;
    }
    const CFold f = folds[jb[1]];
    const CDom L = doms[f.left];
    const CDom R = doms[f.right];
    const int bL = L.off[gp], eL = L.off[gp + 1];
    const int bR = R.off[gp], eR = R.off[gp + 1];
    double uL[3] = {0.0, 0.0, 0.0}, uR[3] = {0.0, 0.0, 0.0};
    for (int k = bL; k < eL; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) uL[c] = uL[c] + u[3 * static_cast<int64_t>(L.ids[k]) + c] * L.sf[k];
    for (int k = bR; k < eR; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) uR[c] = uR[c] + u[3 * static_cast<int64_t>(R.ids[k]) + c] * R.sf[k];
    for (int c = 0; c < 3; ++c) {
        cvals[row0 + c] = uL[c] - uR[c];
        for (int k = bL; k < eL; ++k) jvals[t++] = L.sf[k];
        for (int k = bR; k < eR; ++k) jvals[t++] = -R.sf[k];
    }
    if (!rigid && !f.rigid) return;
    // F0 of both sides (Kinematics.h:75-91) and the fold angle row
    double F[12];
    for (int side = 0; side < 2; ++side) {
        const CDom& D = side ? R : L;
        const int b = side ? bR : bL, e = side ? eR : eL;
        double a1[3] = {1.0, 0.0, 0.0}, a2[3] = {0.0, 1.0, 0.0};
        for (int k = b; k < e; ++k) {
            const double* uk = u + 3 * static_cast<int64_t>(D.ids[k]);
            const double px = D.sf[D.np + k], py = D.sf[2 * D.np + k];
#pragma unroll
            for (int c = 0; c < 3; ++c) { a1[c] = a1[c] + uk[c] * px; a2[c] = a2[c] + uk[c] * py; }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) { F[6 * side + c] = a1[c]; F[6 * side + 3 + c] = a2[c]; }
    }
    double grad[12];
    const double theta = fold_angle_grad(F, L.tx[gp], L.ty[gp], grad);
    cvals[row0 + 3] = theta - (f.theta0 ? f.theta0[gp] : f.Ang0);
    for (int side = 0; side < 2; ++side) {
        const CDom& D = side ? R : L;
        const int b = side ? bR : bL, e = side ? eR : eL;
        const int blk = 6 * side;
        for (int k = b; k < e; ++k) {
            const double dX = D.sf[D.np + k], dY = D.sf[2 * D.np + k];
#pragma unroll
            for (int c = 0; c < 3; ++c) jvals[t++] = grad[blk + c] * dX + grad[blk + c + 3] * dY;
        }
    }
}

// contribution record: dom, pair index, gp, type (0..2 force component, 3 moment);
// value record: coefficient (force component or moment), n_out x, n_out y
__global__ void k_external_force(int ndof, const int* __restrict__ off, const int* __restrict__ rec,
                                 const double* __restrict__ val, const CDom* __restrict__ doms,
                                 double* __restrict__ F) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ndof) return;
    double acc = 0.0;
    for (int r = off[q]; r < off[q + 1]; ++r) {
        const int* rc = rec + 4 * static_cast<int64_t>(r);
        const double* vv = val + 3 * static_cast<int64_t>(r);
        const CDom d = doms[rc[0]];
        const int k = rc[1], g = rc[2];
        if (rc[3] < 3) {
            acc = acc + (d.sf[k] * d.w[g]) * vv[0];
        } else {
            acc = acc + d.w[g] * vv[0] * (d.sf[d.np + k] * vv[1] + d.sf[2 * d.np + k] * vv[2]);
        }
    }
    F[q] = acc;
}

#pragma clang fp contract(on)

namespace {

std::vector<CDom> domain_table(Context& c) {
    std::vector<CDom> t(c.doms.size());
    for (size_t i = 0; i < c.doms.size(); ++i) {
        Domain& d = *c.doms[i];
        CDom& x = t[i];
        x.sf = d.sf.p; x.off = d.off.p; x.ids = d.ids.p; x.np = d.npairs;
        x.nx = d.d_nx.p; x.ny = d.d_ny.p; x.tx = d.d_tx.p; x.ty = d.d_ty.p; x.w = d.d_w.p;
        for (int k = 0; k < 4; ++k) x.val[k] = d.bc.val[k];
        x.active = d.bc.active;
    }
    return t;
}

std::vector<uint64_t> generations(const Context& c) {
    std::vector<uint64_t> g;
    for (const Domain* d : c.doms) g.push_back(d->gen);
    return g;
}

void need_sf(const Domain& d, const char* what) {
    if (d.npairs < 0 || !d.have_sf) throw Error(CVM_E_STATE, std::string(what) + ": shape functions not computed");
}

}  // namespace

void build_constraint_layout(Context& c, bool rigid) {
    ConstraintLayout& L = c.cons;
    L.ready = false;
    L.rows.clear(); L.cols.clear();
    std::vector<int> jobs;
    int row = 0;
    int64_t trip = 0;
    // Dirichlet: point domains, then edge domains, each in creation order
    for (int kind : {CVM_DOMAIN_POINT, CVM_DOMAIN_EDGE})
        for (size_t di = 0; di < c.doms.size(); ++di) {
            Domain& d = *c.doms[di];
            if (d.kind != kind || !(d.bc.active & 15)) continue;
            need_sf(d, "constraint rows");
            for (int g = 0; g < d.ng; ++g) {
                jobs.insert(jobs.end(), {JOB_DIRICHLET, static_cast<int>(di), 0, g, row, static_cast<int>(trip)});
                const int b = d.h_off[g], e = d.h_off[g + 1];
                for (int comp = 0; comp < 4; ++comp) {
                    if (!(d.bc.active & (1 << comp))) continue;
                    const int dof = comp < 3 ? comp : 2;
                    for (int k = b; k < e; ++k) { L.rows.push_back(row); L.cols.push_back(3 * d.h_ids[k] + dof); }
                    trip += e - b;
                    ++row;
                }
            }
        }
    // folds: tie rows (and the angle row when rigid) per active crease GP
    for (size_t fi = 0; fi < c.folds.size(); ++fi) {
        const Fold& f = c.folds[fi];
        Domain& Ld = *c.doms[f.left];
        Domain& Rd = *c.doms[f.right];
        need_sf(Ld, "fold rows"); need_sf(Rd, "fold rows");
        for (int g = 0; g < Ld.ng; ++g) {
            if (!f.active.empty() && !f.active[g]) continue;
            jobs.insert(jobs.end(), {JOB_FOLD, static_cast<int>(fi), 0, g, row, static_cast<int>(trip)});
            const int bL = Ld.h_off[g], eL = Ld.h_off[g + 1], bR = Rd.h_off[g], eR = Rd.h_off[g + 1];
            for (int comp = 0; comp < 3; ++comp) {
                for (int k = bL; k < eL; ++k) { L.rows.push_back(row); L.cols.push_back(3 * Ld.h_ids[k] + comp); }
                for (int k = bR; k < eR; ++k) { L.rows.push_back(row); L.cols.push_back(3 * Rd.h_ids[k] + comp); }
                trip += (eL - bL) + (eR - bR);
                ++row;
            }
            if (rigid || f.rigid) {
                for (int k = bL; k < eL; ++k)
                    for (int comp = 0; comp < 3; ++comp) { L.rows.push_back(row); L.cols.push_back(3 * Ld.h_ids[k] + comp); }
                for (int k = bR; k < eR; ++k)
                    for (int comp = 0; comp < 3; ++comp) { L.rows.push_back(row); L.cols.push_back(3 * Rd.h_ids[k] + comp); }
                trip += 3 * ((eL - bL) + (eR - bR));
                ++row;
            }
        }
    }
    if (trip > 0x7fffffffLL) throw Error(CVM_E_INVALID, "constraint Jacobian has more than 2^31 triplets");
    L.rigid = rigid;
    L.nrows = row;
    L.ntrip = trip;
    L.njobs = static_cast<int>(jobs.size() / JOB_INTS);
    L.d_job.upload(jobs.empty() ? std::vector<int>(JOB_INTS, 0) : jobs, c.stream);
    std::vector<CDom> dt = domain_table(c);
    std::vector<CFold> ft(std::max<size_t>(c.folds.size(), 1));
    for (size_t fi = 0; fi < c.folds.size(); ++fi) {
        const Fold& f = c.folds[fi];
        ft[fi].left = f.left; ft[fi].right = f.right;
        ft[fi].theta0 = f.theta0.empty() ? nullptr : f.d_theta0.p;
        ft[fi].Ang0 = f.Ang0;
        ft[fi].rigid = f.rigid ? 1 : 0;
    }
    L.d_dom.upload(reinterpret_cast<const unsigned char*>(dt.data()), std::max<size_t>(dt.size(), 1) * sizeof(CDom),
                   c.stream);
    L.d_fold.upload(reinterpret_cast<const unsigned char*>(ft.data()), ft.size() * sizeof(CFold), c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    L.gens = generations(c);
    L.ready = true;
}

void launch_constraints(Context& c, const double* xdev, double* cdev, double* jdev) {
    ConstraintLayout& L = c.cons;
    if (!L.ready) throw Error(CVM_E_STATE, "constraint layout not built (cvm_constraint_layout)");
    if (L.gens != generations(c))
        throw Error(CVM_E_STATE, "neighbours or shape functions changed since cvm_constraint_layout");
    if (L.njobs == 0) return;
    const int slot = timing_begin(c, "constraint_rows");
    const int threads = 64;
    hipLaunchKernelGGL(k_constraints, dim3((L.njobs + threads - 1) / threads), dim3(threads), 0, c.stream, L.njobs,
                       L.d_job.p, reinterpret_cast<const CDom*>(L.d_dom.p),
                       reinterpret_cast<const CFold*>(L.d_fold.p), L.rigid ? 1 : 0, xdev, cdev, jdev);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
}

void launch_external_force(Context& c, double* Fdev) {
    ConstraintLayout& L = c.cons;
    const int ndof = 3 * c.nn;
    if (!L.fext_ready || L.fext_gens != generations(c)) {
        // contributions in the reference's order: point forces, then per edge
        // domain its forces and its moments, then fold left / right forces
        struct Rec { int dof, dom, pair, gp, type; double v, nx, ny; };
        std::vector<Rec> recs;
        auto force = [&](int di) {
            Domain& d = *c.doms[di];
            const double* fv = d.bc.force;
            if (fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2] <= 0.0) return;
            need_sf(d, "external force");
            for (int g = 0; g < d.ng; ++g)
                for (int k = d.h_off[g]; k < d.h_off[g + 1]; ++k)
                    for (int comp = 0; comp < 3; ++comp)
                        recs.push_back({3 * d.h_ids[k] + comp, di, k, g, comp, fv[comp], 0.0, 0.0});
        };
        auto moment = [&](int di) {
            Domain& d = *c.doms[di];
            if (d.bc.moment == 0.0) return;
            need_sf(d, "external force");
            for (int g = 0; g < d.ng; ++g) {
                // outward normal: the neighbours lie on the inner side (plate/Meshless.h:84-90)
                double side = 0.0;
                for (int k = d.h_off[g]; k < d.h_off[g + 1]; ++k) {
                    const int v = d.h_ids[k];
                    side += (c.hX[v] - d.hx[g]) * d.hnx[g] + (c.hY[v] - d.hy[g]) * d.hny[g];
                }
                const double nx = side > 0.0 ? -d.hnx[g] : d.hnx[g];
                const double ny = side > 0.0 ? -d.hny[g] : d.hny[g];
                for (int k = d.h_off[g]; k < d.h_off[g + 1]; ++k)
                    recs.push_back({3 * d.h_ids[k] + 2, di, k, g, 3, d.bc.moment, nx, ny});
            }
        };
        for (size_t di = 0; di < c.doms.size(); ++di)
            if (c.doms[di]->kind == CVM_DOMAIN_POINT) force(static_cast<int>(di));
        for (size_t di = 0; di < c.doms.size(); ++di)
            if (c.doms[di]->kind == CVM_DOMAIN_EDGE) { force(static_cast<int>(di)); moment(static_cast<int>(di)); }
        for (const Fold& f : c.folds) { force(f.left); force(f.right); }
        std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.dof < b.dof; });
        std::vector<int> off(ndof + 1, 0), ri;
        std::vector<double> rv;
        for (const Rec& r : recs) {
            off[r.dof + 1]++;
            ri.insert(ri.end(), {r.dom, r.pair, r.gp, r.type});
            rv.insert(rv.end(), {r.v, r.nx, r.ny});
        }
        for (int k = 0; k < ndof; ++k) off[k + 1] += off[k];
        L.d_fx_off.upload(off, c.stream);
        L.d_fx_rec.upload(ri.empty() ? std::vector<int>(4, 0) : ri, c.stream);
        L.d_fx_val.upload(rv.empty() ? std::vector<double>(3, 0.0) : rv, c.stream);
        std::vector<CDom> dt = domain_table(c);
        L.d_dom.upload(reinterpret_cast<const unsigned char*>(dt.data()),
                       std::max<size_t>(dt.size(), 1) * sizeof(CDom), c.stream);
        CVM_HIP(hipStreamSynchronize(c.stream));
        L.fext_gens = generations(c);
        L.fext_ready = true;
        // the descriptor table was refreshed: a constraint layout built on other
        // generations stays invalid through its own generation check
    }
    if (ndof == 0) return;
    const int slot = timing_begin(c, "external_force");
    const int threads = 256;
    hipLaunchKernelGGL(k_external_force, dim3((ndof + threads - 1) / threads), dim3(threads), 0, c.stream, ndof,
                       L.d_fx_off.p, L.d_fx_rec.p, L.d_fx_val.p, reinterpret_cast<const CDom*>(L.d_dom.p), Fdev);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
}

}  // namespace cvm
