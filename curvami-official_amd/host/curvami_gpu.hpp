// C++ host mirror of the reference's hot-path interfaces over the C ABI
// (include/curvami_gpu.h). Header-only; link libcurvami_gpu.so.
//
// Same names, argument meaning and error behaviour as the reference:
//   GlobalFRNN::execute           core/include/curvami/FRNN.h:23-78
//   GlobalRKPM::solve             core/include/curvami/GlobalRKPM.h:19-397
//   NonLinearAssembler            plate/Assembly.h:18-345
//   Meshless (assembly methods)   plate/Meshless.h:30-127
// Errors become std::runtime_error carrying the reference's message text.
// Eigen is not used: sparse K is returned as CSC arrays (outer, inner, values),
// the layout of the reference's Eigen::SparseMatrix<double> (ColMajor, int).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/curvami_gpu.h"

namespace curvami_gpu {

class Context {
public:
    explicit Context(int device = 0) {
        if (cvm_create(device, &ctx_) != CVM_OK) throw std::runtime_error("cvm_create failed (no HIP device?)");
    }
    ~Context() { cvm_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cvm_context* get() const { return ctx_; }
    void check(int st) const {
        if (st != CVM_OK) throw std::runtime_error(cvm_last_error(ctx_));
    }

private:
    cvm_context* ctx_ = nullptr;
};

struct GaussPoints {
    std::vector<double> x, y, w, nx, ny, tx, ty;
    int size() const { return static_cast<int>(x.size()); }
};

struct CSC {
    std::vector<int> outer, inner;
    std::vector<double> values;
};

// GlobalFRNN: one search radius; execute() fills the per-GP neighbour lists
// (CSR: off[ng+1], ids) and s = sqrt(dist2)/rho, in the reference's order.
class GlobalFRNN {
public:
    explicit GlobalFRNN(double radius) : rho_(radius) {}
    void execute(Context& c, int domain, std::vector<int>& off, std::vector<int>& ids, std::vector<double>& s) {
        int64_t np = 0;
        c.check(cvm_neighbours(c.get(), domain, rho_, &np));
        // off must hold ng + 1 entries (ng = the domain's GP count)
        ids.resize(static_cast<size_t>(np));
        s.resize(static_cast<size_t>(np));
        c.check(cvm_get_neighbours(c.get(), domain, off.data(), ids.data(), s.data()));
    }

private:
    double rho_;
};

namespace GlobalRKPM {
// Fills Psi, dPsi, ddPsi of `domain` on the device (reachable through
// Meshless::shapeFunctions). Throws the reference's messages on a singular
// moment matrix or a collapsed pivot.
inline void solve(Context& c, int domain, int order, int kernel = CVM_KERNEL_QUARTIC) {
    int nbad = 0, first = 0;
    c.check(cvm_shape_functions(c.get(), domain, order, kernel, &nbad, &first));
}
}  // namespace GlobalRKPM

// The creased plate with its integration domains and folds; mirrors the parts
// of Meshless that the hot path uses.
class Meshless {
public:
    explicit Meshless(int reproducibility_order, int device = 0) : order(reproducibility_order), ctx_(device) {}

    int order;
    double rho = 0.0;
    int kernel = CVM_KERNEL_QUARTIC;

    void setNodes(const std::vector<double>& X, const std::vector<double>& Y) {
        nn_ = static_cast<int>(X.size());
        ctx_.check(cvm_set_nodes(ctx_.get(), nn_, X.data(), Y.data()));
    }
    // plies: 6 doubles per ply (E1, E2, nu12, G12, angle_deg, thickness)
    void setLaminate(const std::vector<double>& plies) {
        ctx_.check(cvm_set_laminate(ctx_.get(), static_cast<int>(plies.size() / 6), plies.data()));
    }
    // Kinematics::naturalCurvature() (plate/Kinematics.h:20-45): sign -1 mountain, +1 valley
    void setNaturalCurvature(bool active, double x0, double width, double theta0, double sign) {
        ctx_.check(cvm_set_natural_curvature(ctx_.get(), active ? 1 : 0, x0, width, theta0, sign));
    }
    void addCrease(const std::vector<double>& vx, const std::vector<double>& vy, bool ext_s, bool ext_e,
                   bool closed = false) {
        ctx_.check(cvm_add_crease(ctx_.get(), static_cast<int>(vx.size()), vx.data(), vy.data(), ext_s, ext_e, closed));
    }
    int addDomain(int kind, const GaussPoints& g) {
        int id = -1;
        auto ptr = [&](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
        ctx_.check(cvm_add_domain(ctx_.get(), kind, g.size(), g.x.data(), g.y.data(), g.w.data(), ptr(g.nx),
                                  ptr(g.ny), ptr(g.tx), ptr(g.ty), &id));
        sizes_.push_back(g.size());
        return id;
    }
    // Meshless::addFold (Meshless.h:30-46): the caller passes the already
    // offset left/right crease GPs (X_ref +- 1e-7 n).
    int addFold(const GaussPoints& left, const GaussPoints& right, double Kt, double Ang0) {
        const int l = addDomain(CVM_DOMAIN_FOLD_LEFT, left);
        const int r = addDomain(CVM_DOMAIN_FOLD_RIGHT, right);
        int id = -1;
        ctx_.check(cvm_add_fold(ctx_.get(), l, r, Kt, Ang0, &id));
        return id;
    }
    // SetupHelper::recomputeShapeFunctions (SetupHelper.h:112-138): FRNN + RKPM
    // on every domain. The first call (and any after a domain was added or the
    // radius changed) searches and solves domain by domain; the others search
    // again on the device and solve every domain in one call
    // (cvm_recompute_shape_functions), with the same results.
    void recomputeShapeFunctions() {
        cons_ready_ = false;
        if (searched_ > 0 && searched_ == static_cast<int>(sizes_.size()) && searched_rho_ == rho) {
            int changed = 0, failed = -1, nbad = 0, first = 0;
            ctx_.check(cvm_recompute_shape_functions(ctx_.get(), order, kernel, &changed, &failed, &nbad, &first));
            return;
        }
        for (int d = 0; d < static_cast<int>(sizes_.size()); ++d) {
            GlobalFRNN frnn(rho);
            std::vector<int> off(static_cast<size_t>(sizes_[d]) + 1), ids;
            std::vector<double> s;
            frnn.execute(ctx_, d, off, ids, s);
            GlobalRKPM::solve(ctx_, d, order, kernel);
        }
        searched_ = static_cast<int>(sizes_.size());
        searched_rho_ = rho;
    }
    // Meshless::initAssembly -> NonLinearAssembler::initSparsityPattern
    void initAssembly() { ctx_.check(cvm_init_assembly(ctx_.get(), &nnz_face_, &nnz_union_)); }

    // Meshless::assembleSystem (Meshless.h:113-127)
    void assembleSystem(CSC& K, std::vector<double>& F_int, const std::vector<double>& x, bool flagRig_current) {
        const bool folds = !flagRig_current;
        K.outer.resize(3 * static_cast<size_t>(nn_) + 1);
        K.inner.resize(static_cast<size_t>(folds ? nnz_union_ : nnz_face_));
        K.values.resize(K.inner.size());
        ctx_.check(cvm_get_pattern(ctx_.get(), folds, K.outer.data(), K.inner.data()));
        F_int.resize(3 * static_cast<size_t>(nn_));
        const int flags = CVM_RESIDUAL | CVM_TANGENT | (folds ? CVM_ELASTIC_FOLDS : 0);
        ctx_.check(cvm_assemble(ctx_.get(), x.data(), flags, F_int.data(), K.values.data()));
    }
    // Meshless::assembleSystemInternalOnly (Meshless.h:102-111)
    std::vector<double> assembleSystemInternalOnly(const std::vector<double>& x, bool flagRig_current) {
        std::vector<double> F(3 * static_cast<size_t>(nn_));
        const int flags = CVM_RESIDUAL | (flagRig_current ? 0 : CVM_ELASTIC_FOLDS);
        ctx_.check(cvm_assemble(ctx_.get(), x.data(), flags, F.data(), nullptr));
        return F;
    }
    // BoundaryCondition of a domain (PlateData.h:36-50); active bit 0 u, 1 v, 2 w, 3 rotation.
    void setBoundaryCondition(int domain, int active, const double val[4], const double force[3], double moment) {
        ctx_.check(cvm_set_domain_bc(ctx_.get(), domain, active, val, force, moment));
        cons_ready_ = false;
    }
    // Meshless::evaluateConstraints (Meshless.h:129-195): c and J as triplets in the
    // reference's push order (J_all.setFromTriplets sums duplicates).
    struct Triplet { int row, col; double value; };
    void evaluateConstraints(const std::vector<double>& x, std::vector<double>& c_all,
                             std::vector<Triplet>& J_triplets, int& nrows, bool flagRig_current) {
        if (!cons_ready_ || cons_rigid_ != flagRig_current) {
            int64_t nr = 0, nt = 0;
            ctx_.check(cvm_constraint_layout(ctx_.get(), flagRig_current, &nr, &nt));
            crow_.resize(static_cast<size_t>(nt));
            ccol_.resize(static_cast<size_t>(nt));
            ctx_.check(cvm_get_constraint_pattern(ctx_.get(), crow_.data(), ccol_.data()));
            cnrows_ = static_cast<int>(nr);
            cons_ready_ = true;
            cons_rigid_ = flagRig_current;
        }
        c_all.resize(static_cast<size_t>(cnrows_));
        std::vector<double> v(crow_.size());
        ctx_.check(cvm_constraints(ctx_.get(), x.data(), 0, c_all.data(), v.data()));
        J_triplets.resize(v.size());
        for (size_t k = 0; k < v.size(); ++k) J_triplets[k] = {crow_[k], ccol_[k], v[k]};
        nrows = cnrows_;
    }
    // Meshless::computeExternalForce (Meshless.h:60-100)
    std::vector<double> computeExternalForce() {
        std::vector<double> F(3 * static_cast<size_t>(nn_));
        ctx_.check(cvm_external_force(ctx_.get(), 0, F.data()));
        return F;
    }
    // NonLinearAssembler::csrMapping (Assembly.h:54)
    std::vector<int> csrMapping(int gp) {
        int n = 0;
        ctx_.check(cvm_csr_mapping(ctx_.get(), gp, nullptr, 0, &n));
        std::vector<int> out(static_cast<size_t>(n));
        ctx_.check(cvm_csr_mapping(ctx_.get(), gp, out.data(), n, &n));
        return out;
    }
    Context& context() { return ctx_; }

private:
    Context ctx_;
    int nn_ = 0;
    int64_t nnz_face_ = 0, nnz_union_ = 0;
    std::vector<int> sizes_;
    int searched_ = 0;            // domains searched by the last full recomputeShapeFunctions
    double searched_rho_ = 0.0;   // and its radius
    bool cons_ready_ = false, cons_rigid_ = false;
    int cnrows_ = 0;
    std::vector<int> crow_, ccol_;
};

}  // namespace curvami_gpu
