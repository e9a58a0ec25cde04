// C ABI of the curvami GPU library: argument checks, host orchestration,
// status codes. Reference interfaces replaced are cited in include/curvami_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "cvm_internal.h"

namespace cvm {

Context::~Context() {
    for (Domain* d : doms) delete d;
    doms.clear();
    for (size_t i = 0; i < timing.ev0.size(); ++i) {
        (void)hipEventDestroy(timing.ev0[i]);
        (void)hipEventDestroy(timing.ev1[i]);
    }
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (side) (void)hipStreamSynchronize(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
}

hipStream_t side_stream(Context& c) {
    if (!c.side) {
        CVM_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
        CVM_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
        CVM_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    }
    return c.side;
}

// ---- per-kernel event timing (the context's stream, or its side stream) ---------
static int kernel_id(Timing& t, const char* name) {
    for (size_t i = 0; i < t.names.size(); ++i)
        if (t.names[i] == name) return static_cast<int>(i);
    t.names.emplace_back(name);
    t.total_ms.push_back(0.0);
    t.count.push_back(0);
    return static_cast<int>(t.names.size()) - 1;
}

static void timing_collect(Context& c) {
    Timing& t = c.timing;
    if (t.pending.empty()) return;
    CVM_HIP(hipStreamSynchronize(c.stream));
    if (c.side) CVM_HIP(hipStreamSynchronize(c.side));   // kernels timed on the side stream
    for (size_t i = 0; i < t.pending.size(); ++i) {
        float ms = 0.0f;
        CVM_HIP(hipEventElapsedTime(&ms, t.ev0[i], t.ev1[i]));
        t.total_ms[t.pending[i]] += ms;
        t.count[t.pending[i]] += 1;
    }
    t.pending.clear();
}

int timing_begin(Context& c, const char* name, hipStream_t st) {
    Timing& t = c.timing;
    if (!t.on) return -1;
    const int id = kernel_id(t, name);
    const size_t slot = t.pending.size();
    if (slot >= 4096) timing_collect(c);
    const size_t s2 = t.pending.size();
    while (t.ev0.size() <= s2) {
        hipEvent_t a, b;
        CVM_HIP(hipEventCreate(&a));
        CVM_HIP(hipEventCreate(&b));
        t.ev0.push_back(a);
        t.ev1.push_back(b);
    }
    t.pending.push_back(id);
    CVM_HIP(hipEventRecord(t.ev0[s2], st ? st : c.stream));
    return static_cast<int>(s2);
}

void timing_end(Context& c, int slot, hipStream_t st) {
    if (slot < 0) return;
    CVM_HIP(hipEventRecord(c.timing.ev1[slot], st ? st : c.stream));
}

void sync_nodes(Context& c) {
    if (!c.nodes_dirty) return;
    c.X.upload(c.hX, c.stream);
    c.Y.upload(c.hY, c.stream);
    if (c.batched) c.d_nsheet.upload(c.h_nsheet, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    c.nodes_dirty = false;
}

void sync_domain(Context& c, Domain& d) {
    if (!d.dirty) return;
    d.gx.upload(d.hx, c.stream);
    d.gy.upload(d.hy, c.stream);
    d.d_w.upload(d.hw, c.stream);
    d.d_tx.upload(d.htx, c.stream);
    d.d_ty.upload(d.hty, c.stream);
    d.d_nx.upload(d.hnx, c.stream);
    d.d_ny.upload(d.hny, c.stream);
    if (c.batched) d.d_gsheet.upload(d.h_gsheet, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    d.dirty = false;
}

static Domain& domain(Context& c, int id) {
    if (id < 0 || id >= static_cast<int>(c.doms.size())) throw Error(CVM_E_INVALID, "bad domain id");
    return *c.doms[id];
}

static void laminate_abd(const double* plies, int nply, double* A, double* B, double* D) {
    // Constitutive::computeABD (plate/Constitutive.h:10-50)
    for (int k = 0; k < 9; ++k) { A[k] = 0.0; B[k] = 0.0; D[k] = 0.0; }
    double h = 0.0;
    for (int i = 0; i < nply; ++i) h += plies[6 * i + 5];
    double z0 = -h / 2.0;
    for (int i = 0; i < nply; ++i) {
        const double E1 = plies[6 * i], E2 = plies[6 * i + 1], nu12 = plies[6 * i + 2], G12 = plies[6 * i + 3];
        const double ang = plies[6 * i + 4], th = plies[6 * i + 5];
        const double nu21 = (E2 / E1) * nu12;
        const double den = 1.0 - nu12 * nu21;
        const double Q00 = E1 / den, Q11 = E2 / den, Q01 = (nu12 * E2) / den, Q22 = G12;
        const double t = ang * M_PI / 180.0;
        const double m = std::cos(t), n = std::sin(t);
        const double m2 = m * m, n2 = n * n, m4 = m2 * m2, n4 = n2 * n2;
        double Qb[9];
        Qb[0] = Q00 * m4 + 2.0 * (Q01 + 2.0 * Q22) * m2 * n2 + Q11 * n4;
        Qb[4] = Q00 * n4 + 2.0 * (Q01 + 2.0 * Q22) * m2 * n2 + Q11 * m4;
        Qb[1] = Qb[3] = (Q00 + Q11 - 4.0 * Q22) * m2 * n2 + Q01 * (m4 + n4);
        Qb[2] = Qb[6] = (Q00 - Q01 - 2.0 * Q22) * n * m * m2 + (Q01 - Q11 + 2.0 * Q22) * n2 * n * m;
        Qb[5] = Qb[7] = (Q00 - Q01 - 2.0 * Q22) * m * n * n2 + (Q01 - Q11 + 2.0 * Q22) * m2 * m * n;
        Qb[8] = (Q00 + Q11 - 2.0 * Q01 - 2.0 * Q22) * m2 * n2 + Q22 * (m4 + n4);
        const double z1 = z0 + th;
        for (int k = 0; k < 9; ++k) {
            A[k] += Qb[k] * (z1 - z0);
            B[k] += 0.5 * Qb[k] * (z1 * z1 - z0 * z0);
            D[k] += (1.0 / 3.0) * Qb[k] * (z1 * z1 * z1 - z0 * z0 * z0);
        }
        z0 = z1;
    }
}

}  // namespace cvm

using cvm::Context;
using cvm::Error;

#define CVM_TRY(ctx)                                              \
    if (!(ctx)) return CVM_E_INVALID;                             \
    try {
#define CVM_CATCH(ctx)                                            \
    }                                                             \
    catch (const cvm::Error& e) {                                 \
        (ctx)->c.err = e.what();                                  \
        return e.code;                                            \
    }                                                             \
    catch (const std::bad_alloc&) {                               \
        (ctx)->c.err = "host allocation failed";                  \
        return CVM_E_NO_MEMORY;                                   \
    }                                                             \
    catch (const std::exception& e) {                             \
        (ctx)->c.err = e.what();                                  \
        return CVM_E_INVALID;                                     \
    }                                                             \
    return CVM_OK;

extern "C" {

int cvm_create(int device, cvm_context** out) {
    if (!out) return CVM_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CVM_E_HIP;
    if (device < 0 || device >= ndev) return CVM_E_INVALID;
    cvm_context* ctx = new (std::nothrow) cvm_context();
    if (!ctx) return CVM_E_NO_MEMORY;
    ctx->c.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CVM_E_HIP;
    }
    ctx->c.ws.stream = ctx->c.stream;
    ctx->c.ws.timed = true;
    cvm::noalloc_scratch_init(device);   // device address of this GPU's static scratch (device current)
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        ctx->c.lds_limit = static_cast<size_t>(lds);
    *out = ctx;
    return CVM_OK;
}

void cvm_destroy(cvm_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
}

const char* cvm_last_error(const cvm_context* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int cvm_set_nodes(cvm_context* ctx, int nn, const double* X, const double* Y) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (nn < 0 || (nn > 0 && (!X || !Y))) throw Error(CVM_E_INVALID, "bad node arrays");
    if (c.batched) throw Error(CVM_E_STATE, "batched context: nodes come with cvm_add_sheet");
    CVM_HIP(hipSetDevice(c.device));
    c.nn = nn;
    c.hX.assign(X, X + nn);
    c.hY.assign(Y, Y + nn);
    c.X.upload(c.hX, c.stream);
    c.Y.upload(c.hY, c.stream);
    c.plan.ready = false;
    c.node_sign_creases = -1;
    for (cvm::Domain* d : c.doms) { d->npairs = -1; d->have_sf = false; }
    CVM_HIP(hipStreamSynchronize(c.stream));
    CVM_CATCH(ctx)
}

int cvm_set_laminate(cvm_context* ctx, int nply, const double* plies) {
    CVM_TRY(ctx)
    if (nply <= 0 || !plies) throw Error(CVM_E_INVALID, "need at least one ply");
    cvm::laminate_abd(plies, nply, ctx->c.A, ctx->c.B, ctx->c.D);
    ctx->c.have_mat = true;
    CVM_CATCH(ctx)
}

int cvm_set_natural_curvature(cvm_context* ctx, int active, double x0, double width, double theta0, double sign) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (active && !(width > 0.0)) throw Error(CVM_E_INVALID, "natural_curvature.width must be > 0");
    if (active && sign != 1.0 && sign != -1.0) throw Error(CVM_E_INVALID, "natural curvature sign must be +1 or -1");
    CVM_HIP(hipSetDevice(c.device));
    c.nat.active = active != 0;
    c.nat.x0 = x0; c.nat.width = width; c.nat.theta0 = theta0; c.nat.sign = sign;
    c.kbar_ng = -1;
    if (c.plan.ready) (void)cvm::natural_kappa(c, true);
    CVM_CATCH(ctx)
}

int cvm_get_abd(const cvm_context* ctx, double* A, double* B, double* D) {
    if (!ctx) return CVM_E_INVALID;
    for (int k = 0; k < 9; ++k) {
        if (A) A[k] = ctx->c.A[k];
        if (B) B[k] = ctx->c.B[k];
        if (D) D[k] = ctx->c.D[k];
    }
    return CVM_OK;
}

int cvm_add_crease(cvm_context* ctx, int nv, const double* vx, const double* vy, int extend_start, int extend_end,
                   int is_closed) {
    CVM_TRY(ctx)
    if (nv < 0 || (nv > 0 && (!vx || !vy))) throw Error(CVM_E_INVALID, "bad crease arrays");
    if (ctx->c.batched) {
        if (ctx->c.sheets.empty()) throw Error(CVM_E_STATE, "batched context: add a sheet first");
        ctx->c.sheets.back().ncreases++;      // creases belong to the last sheet added
    }
    cvm::Crease cr;
    cr.vx.assign(vx, vx + nv);
    cr.vy.assign(vy, vy + nv);
    cr.ext_s = extend_start ? 1 : 0;
    cr.ext_e = extend_end ? 1 : 0;
    cr.closed = is_closed ? 1 : 0;
    ctx->c.creases.push_back(cr);
    ctx->c.creases_dirty = true;
    ctx->c.node_sign_creases = -1;
    for (cvm::Domain* d : ctx->c.doms) { d->have_sf = false; d->ind_creases = -1; }
    CVM_CATCH(ctx)
}

int cvm_add_domain(cvm_context* ctx, int kind, int ng, const double* gx, const double* gy, const double* w,
                   const double* nx, const double* ny, const double* tx, const double* ty, int* domain_id) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (kind < CVM_DOMAIN_FACE || kind > CVM_DOMAIN_FOLD_RIGHT) throw Error(CVM_E_INVALID, "bad domain kind");
    if (ng < 0 || (ng > 0 && (!gx || !gy || !w))) throw Error(CVM_E_INVALID, "bad Gauss-point arrays");
    CVM_HIP(hipSetDevice(c.device));
    auto append = [&](cvm::Domain& d) {
        d.hx.insert(d.hx.end(), gx, gx + ng);
        d.hy.insert(d.hy.end(), gy, gy + ng);
        d.hw.insert(d.hw.end(), w, w + ng);
        auto add = [&](std::vector<double>& v, const double* a) {
            if (a) v.insert(v.end(), a, a + ng); else v.insert(v.end(), static_cast<size_t>(ng), 0.0);
        };
        add(d.hnx, nx); add(d.hny, ny); add(d.htx, tx); add(d.hty, ty);
        if (c.batched) d.h_gsheet.insert(d.h_gsheet.end(), static_cast<size_t>(ng), static_cast<int>(c.sheets.size()) - 1);
        d.ng += ng;
        d.dirty = true;
        d.npairs = -1;
        d.have_sf = false;
        d.ind_creases = -1;
    };
    if (c.batched && c.sheets.empty()) throw Error(CVM_E_STATE, "batched context: add a sheet first");
    int id = -1;
    if (c.batched && kind == CVM_DOMAIN_FACE) {
        // one face domain holds the face GPs of every sheet, in sheet order
        for (size_t i = 0; i < c.doms.size(); ++i)
            if (c.doms[i]->kind == CVM_DOMAIN_FACE) id = static_cast<int>(i);
    }
    if (id >= 0) {
        append(*c.doms[id]);
    } else {
        auto* d = new cvm::Domain();
        d->kind = kind;
        append(*d);
        c.doms.push_back(d);
        id = static_cast<int>(c.doms.size()) - 1;
    }
    if (!c.batched) cvm::sync_domain(c, *c.doms[id]);
    if (kind == CVM_DOMAIN_FACE) c.plan.ready = false;
    if (domain_id) *domain_id = id;
    CVM_CATCH(ctx)
}

int cvm_add_sheet(cvm_context* ctx, int nn, const double* X, const double* Y, double rho, int* sheet_id,
                  int* node0) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (nn <= 0 || !X || !Y) throw Error(CVM_E_INVALID, "bad node arrays");
    if (!(rho > 0.0)) throw Error(CVM_E_INVALID, "rho must be > 0");
    if (!c.batched && (c.nn > 0 || !c.creases.empty() || !c.doms.empty()))
        throw Error(CVM_E_STATE, "cvm_add_sheet on a context already set up with cvm_set_nodes");
    c.batched = true;
    cvm::SheetSeg sg;
    sg.node0 = c.nn;
    sg.nn = nn;
    sg.crease0 = static_cast<int>(c.creases.size());
    sg.ncreases = 0;
    sg.rho = rho;
    c.sheets.push_back(sg);
    c.hX.insert(c.hX.end(), X, X + nn);
    c.hY.insert(c.hY.end(), Y, Y + nn);
    c.h_nsheet.insert(c.h_nsheet.end(), static_cast<size_t>(nn), static_cast<int>(c.sheets.size()) - 1);
    c.nn += nn;
    c.nodes_dirty = true;
    c.plan.ready = false;
    c.node_sign_creases = -1;
    for (cvm::Domain* d : c.doms) { d->npairs = -1; d->have_sf = false; d->ind_creases = -1; }
    if (sheet_id) *sheet_id = static_cast<int>(c.sheets.size()) - 1;
    if (node0) *node0 = sg.node0;
    CVM_CATCH(ctx)
}

int cvm_add_fold(cvm_context* ctx, int left, int right, double Kt, double Ang0, int* fold_id) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    cvm::Domain& L = cvm::domain(c, left);
    cvm::Domain& R = cvm::domain(c, right);
    if (L.ng != R.ng) throw Error(CVM_E_INVALID, "fold left/right domains differ in size");
    cvm::Fold f;
    f.left = left; f.right = right; f.Kt = Kt; f.Ang0 = Ang0;
    c.folds.push_back(std::move(f));
    c.plan.ready = false;
    if (fold_id) *fold_id = static_cast<int>(c.folds.size()) - 1;
    CVM_CATCH(ctx)
}

int cvm_set_fold_gp_state(cvm_context* ctx, int fold_id, int n, const double* theta0, const char* active) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (fold_id < 0 || fold_id >= static_cast<int>(c.folds.size())) throw Error(CVM_E_INVALID, "bad fold id");
    cvm::Fold& f = c.folds[fold_id];
    if ((theta0 || active) && n != c.doms[f.left]->ng) throw Error(CVM_E_INVALID, "state size differs from crease GP count");
    const bool active_changed = (active != nullptr) != !f.active.empty() ||
                                (active && !std::equal(active, active + n, f.active.begin()));
    f.theta0.assign(theta0, theta0 + (theta0 ? n : 0));
    f.active.assign(active, active + (active ? n : 0));
    if (theta0) f.d_theta0.upload(f.theta0, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    if (active_changed) c.plan.ready = false;   // the fold pattern depends on activity
    c.cons.ready = false;                        // so do the fold rows
    if (c.plan.ready) cvm::upload_fold_descriptors(c);   // theta0 pointer
    CVM_CATCH(ctx)
}

int cvm_set_fold_rigid(cvm_context* ctx, int fold_id, int rigid) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (fold_id < 0 || fold_id >= static_cast<int>(c.folds.size())) throw Error(CVM_E_INVALID, "bad fold id");
    if (c.folds[fold_id].rigid != (rigid != 0)) {
        c.folds[fold_id].rigid = rigid != 0;
        c.plan.ready = false;      // the hinge blocks of the pattern change
        c.cons.ready = false;      // and so do the angle rows
    }
    CVM_CATCH(ctx)
}

int cvm_neighbours(cvm_context* ctx, int domain, double rho, int64_t* npairs) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::Domain& d = cvm::domain(c, domain);
    cvm::launch_frnn(c, d, rho);
    ++d.gen;
    if (npairs) *npairs = d.npairs;
    if (d.kind == CVM_DOMAIN_FACE || d.kind == CVM_DOMAIN_FOLD_LEFT || d.kind == CVM_DOMAIN_FOLD_RIGHT)
        c.plan.ready = false;
    CVM_CATCH(ctx)
}

int cvm_refresh_neighbours(cvm_context* ctx, int* changed) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    bool any = false, pattern = false;
    const std::vector<char> ch = cvm::refresh_frnn_all(c);
    for (size_t k = 0; k < c.doms.size(); ++k) {
        cvm::Domain* d = c.doms[k];
        if (!ch[k]) continue;
        any = true;
        ++d->gen;
        if (d->kind == CVM_DOMAIN_FACE || d->kind == CVM_DOMAIN_FOLD_LEFT || d->kind == CVM_DOMAIN_FOLD_RIGHT)
            pattern = true;
    }
    if (pattern) c.plan.ready = false;
    if (changed) *changed = any ? 1 : 0;
    CVM_CATCH(ctx)
}

int cvm_get_neighbours(cvm_context* ctx, int domain, int* off, int* ids, double* s) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    cvm::Domain& d = cvm::domain(c, domain);
    if (d.npairs < 0) throw Error(CVM_E_STATE, "neighbours not computed");
    if (off) std::copy(d.h_off.begin(), d.h_off.end(), off);
    if (ids) std::copy(d.h_ids.begin(), d.h_ids.end(), ids);
    if (s && d.npairs) {
        d.s.download(s, d.npairs, c.stream);
        CVM_HIP(hipStreamSynchronize(c.stream));
    }
    CVM_CATCH(ctx)
}

static void rkpm_error(const cvm::Domain& d, int code, int nb, int fb, const std::string& where) {
    const int nfirst = d.h_off[fb + 1] - d.h_off[fb];
    if (code == CVM_E_SINGULAR_MOMENT)
        throw Error(code, where + "RKPM: singular moment matrix at " + std::to_string(nb) + " of " + std::to_string(d.ng) +
                              " Gauss points (first is index " + std::to_string(fb) + " with " +
                              std::to_string(nfirst) +
                              " neighbours). The support is rank-deficient for the requested basis order - "
                              "typically too few visible nodes, which happens where a crease cuts most of a "
                              "support away. Increase the support radius, refine the mesh there, or lower the order.");
    throw Error(code, where + "RKPM: moment-matrix pivot collapsed during block inversion at " + std::to_string(nb) +
                          " of " + std::to_string(d.ng) + " Gauss points (first is index " + std::to_string(fb) +
                          " with " + std::to_string(nfirst) +
                          " neighbours). Same cause and remedies as a singular moment matrix.");
}

int cvm_shape_functions(cvm_context* ctx, int domain, int order, int kernel, int* nbad, int* first_bad) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::Domain& d = cvm::domain(c, domain);
    int code = 0, nb = 0, fb = 0;
    cvm::launch_rkpm(c, d, d.rho, order, kernel, &code, &nb, &fb);
    ++d.gen;
    if (nbad) *nbad = nb;
    if (first_bad) *first_bad = fb;
    if (code != CVM_OK) rkpm_error(d, code, nb, fb, "");
    CVM_CATCH(ctx)
}

int cvm_shape_functions_all(cvm_context* ctx, int order, int kernel, int* failed_domain, int* nbad, int* first_bad) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    int dom = -1, code = 0, nb = 0, fb = 0;
    cvm::launch_rkpm_all(c, order, kernel, &dom, &code, &nb, &fb);
    for (cvm::Domain* d : c.doms) ++d->gen;
    if (failed_domain) *failed_domain = dom;
    if (nbad) *nbad = nb;
    if (first_bad) *first_bad = fb;
    if (code != CVM_OK) rkpm_error(*c.doms[dom], code, nb, fb, "domain " + std::to_string(dom) + ": ");
    CVM_CATCH(ctx)
}

int cvm_recompute_shape_functions(cvm_context* ctx, int order, int kernel, int* changed, int* failed_domain,
                                  int* nbad, int* first_bad) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (order < 1 || order > 3) throw Error(CVM_E_INVALID, "order must be 1, 2 or 3");
    if (kernel != CVM_KERNEL_QUARTIC && kernel != CVM_KERNEL_CUBIC_BSPLINE)
        throw Error(CVM_E_INVALID, "unknown window kernel");
    for (cvm::Domain* d : c.doms)
        if (d->npairs < 0) throw Error(CVM_E_STATE, "neighbours not computed for this domain");
    // the shape functions of every domain beside the verification, from the
    // resident lists; formed again from the new lists if any list changed
    bool enq = false;
    cvm::RkpmBatch rb;
    std::vector<int> hs(4 * std::max<size_t>(c.doms.size(), 1));
    const std::vector<char> ch = cvm::refresh_frnn_all(
        c,
        [&](hipStream_t sd) {
            rb = cvm::enqueue_rkpm_all(c, order, kernel, sd);
            enq = true;
        },
        [&] {
            if (enq) cvm::read_rkpm_status(c, hs.data());   // read back with the verification flags
        });
    bool any = false, pattern = false;
    for (size_t k = 0; k < c.doms.size(); ++k) {
        cvm::Domain* d = c.doms[k];
        if (!ch[k]) continue;
        any = true;
        ++d->gen;
        if (d->kind == CVM_DOMAIN_FACE || d->kind == CVM_DOMAIN_FOLD_LEFT || d->kind == CVM_DOMAIN_FOLD_RIGHT)
            pattern = true;
    }
    if (pattern) c.plan.ready = false;
    int dom = -1, code = 0, nb = 0, fb = 0;
    if (enq && !any) cvm::finish_rkpm_all(c, rb, hs.data(), &dom, &code, &nb, &fb);
    else cvm::launch_rkpm_all(c, order, kernel, &dom, &code, &nb, &fb);
    for (cvm::Domain* d : c.doms) ++d->gen;
    if (changed) *changed = any ? 1 : 0;
    if (failed_domain) *failed_domain = dom;
    if (nbad) *nbad = nb;
    if (first_bad) *first_bad = fb;
    if (code != CVM_OK) rkpm_error(*c.doms[dom], code, nb, fb, "domain " + std::to_string(dom) + ": ");
    CVM_CATCH(ctx)
}

int cvm_get_shape_functions(cvm_context* ctx, int domain, double* psi, double* dpx, double* dpy, double* dxx,
                            double* dyy, double* dxy) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    cvm::Domain& d = cvm::domain(c, domain);
    if (!d.have_sf) throw Error(CVM_E_STATE, "shape functions not computed");
    double* outs[6] = {psi, dpx, dpy, dxx, dyy, dxy};
    for (int k = 0; k < 6; ++k)
        if (outs[k] && d.npairs)
            CVM_HIP(hipMemcpyAsync(outs[k], d.sf.p + k * d.npairs, d.npairs * sizeof(double), hipMemcpyDeviceToHost,
                                   c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    CVM_CATCH(ctx)
}

int cvm_init_assembly(cvm_context* ctx, int64_t* nnz_face, int64_t* nnz_union) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (!c.have_mat) throw Error(CVM_E_STATE, "laminate not set");
    cvm::build_plan(c);
    c.kbar_ng = -1;
    (void)cvm::natural_kappa(c, true);
    if (nnz_face) *nnz_face = static_cast<int64_t>(c.plan.face_inner.size());
    if (nnz_union) *nnz_union = static_cast<int64_t>(c.plan.uni_inner.size());
    CVM_CATCH(ctx)
}

int cvm_get_pattern(cvm_context* ctx, int with_folds, int* outer, int* inner) {
    CVM_TRY(ctx)
    const cvm::Plan& P = ctx->c.plan;
    if (!P.ready) throw Error(CVM_E_STATE, "assembly plan not initialised");
    const auto& o = with_folds ? P.uni_outer : P.face_outer;
    const auto& i = with_folds ? P.uni_inner : P.face_inner;
    if (outer) std::copy(o.begin(), o.end(), outer);
    if (inner) std::copy(i.begin(), i.end(), inner);
    CVM_CATCH(ctx)
}

int cvm_csr_mapping(cvm_context* ctx, int gp, int* out, int cap, int* n) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    const cvm::Plan& P = c.plan;
    if (!P.ready) throw Error(CVM_E_STATE, "assembly plan not initialised");
    cvm::Domain* fd = nullptr;
    for (cvm::Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) fd = d;
    if (gp < 0 || gp >= fd->ng) throw Error(CVM_E_INVALID, "bad GP index");
    const int b = fd->h_off[gp], m = fd->h_off[gp + 1] - b;
    const int cnt = m * (m + 1) / 2;
    if (n) *n = cnt;
    if (!out) return CVM_OK;
    if (cap < cnt) throw Error(CVM_E_INVALID, "output too small");
    int mo = 0;
    for (int i = 0; i < m; ++i)
        for (int j = i; j < m; ++j) {
            const int ni = fd->h_ids[b + i], nj = fd->h_ids[b + j];
            const int lo = std::min(ni, nj), hi = std::max(ni, nj);
            const auto& ch = P.face_rows[hi];
            auto it = std::lower_bound(ch.begin(), ch.end(), lo);
            out[mo++] = (it == ch.end() || *it != lo) ? -1 : static_cast<int>(it - ch.begin());
        }
    CVM_CATCH(ctx)
}

int cvm_assemble(cvm_context* ctx, const double* x, int flags, double* F_int, double* K_values) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::Plan& P = c.plan;
    if (!P.ready) throw Error(CVM_E_STATE, "assembly plan not initialised (cvm_init_assembly)");
    if (!x) throw Error(CVM_E_INVALID, "x is null");
    for (cvm::Domain* d : c.doms)
        if ((d->kind == CVM_DOMAIN_FACE) && !d->have_sf) throw Error(CVM_E_STATE, "face shape functions not computed");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    const bool tangent = (flags & CVM_TANGENT) != 0;
    const int64_t nnz = (flags & CVM_ELASTIC_FOLDS) ? static_cast<int64_t>(P.uni_inner.size())
                                                    : static_cast<int64_t>(P.face_inner.size());
    const double* xd = x;
    if (!dev) {
        c.d_x.upload(x, 3 * static_cast<size_t>(c.nn), c.stream);
        xd = c.d_x.p;
    }
    double* Fd = F_int;
    double* Kd = K_values;
    if (!dev) {
        c.d_F.alloc(std::max<size_t>(3 * static_cast<size_t>(c.nn), 1));
        Fd = c.d_F.p;
        if (tangent) {
            c.d_K.alloc(std::max<int64_t>(nnz, 1));
            Kd = c.d_K.p;
        }
    } else if (!Fd || (tangent && !Kd)) {
        throw Error(CVM_E_INVALID, "device output pointer is null");
    }
    cvm::launch_assembly(c, c.ws, flags, xd, Fd, Kd);
    if (!dev) {
        if (F_int) CVM_HIP(hipMemcpyAsync(F_int, Fd, 3 * sizeof(double) * c.nn, hipMemcpyDeviceToHost, c.stream));
        if (tangent && K_values && nnz)
            CVM_HIP(hipMemcpyAsync(K_values, Kd, sizeof(double) * nnz, hipMemcpyDeviceToHost, c.stream));
        CVM_HIP(hipStreamSynchronize(c.stream));
    }
    CVM_CATCH(ctx)
}

int cvm_assemble_noalloc(cvm_context* ctx, const double* x_dev, int flags, double* F_dev, double* K_dev) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::Plan& P = c.plan;
    if (!P.ready) throw Error(CVM_E_STATE, "assembly plan not initialised (cvm_init_assembly)");
    for (cvm::Domain* d : c.doms)
        if ((d->kind == CVM_DOMAIN_FACE) && !d->have_sf) throw Error(CVM_E_STATE, "face shape functions not computed");
    if (!x_dev || !F_dev || ((flags & CVM_TANGENT) && !K_dev)) throw Error(CVM_E_INVALID, "null device pointer");
    cvm::launch_assembly_noalloc(c, flags & ~CVM_DEVICE_PTRS, x_dev, F_dev, K_dev);
    CVM_CATCH(ctx)
}

int cvm_set_domain_bc(cvm_context* ctx, int domain, int active, const double* val, const double* force,
                      double moment) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    cvm::Domain& d = cvm::domain(c, domain);
    if (active & ~15) throw Error(CVM_E_INVALID, "bad active mask (bits 0..3: u, v, w, rotation)");
    d.bc.active = active;
    for (int k = 0; k < 4; ++k) d.bc.val[k] = val ? val[k] : 0.0;
    for (int k = 0; k < 3; ++k) d.bc.force[k] = force ? force[k] : 0.0;
    d.bc.moment = moment;
    c.cons.ready = false;
    c.cons.fext_ready = false;
    CVM_CATCH(ctx)
}

int cvm_constraint_layout(cvm_context* ctx, int rigid, int64_t* nrows, int64_t* ntriplets) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::build_constraint_layout(c, rigid != 0);
    if (nrows) *nrows = c.cons.nrows;
    if (ntriplets) *ntriplets = c.cons.ntrip;
    CVM_CATCH(ctx)
}

int cvm_get_constraint_pattern(cvm_context* ctx, int* rows, int* cols) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (!c.cons.ready) throw Error(CVM_E_STATE, "constraint layout not built (cvm_constraint_layout)");
    if (rows) std::copy(c.cons.rows.begin(), c.cons.rows.end(), rows);
    if (cols) std::copy(c.cons.cols.begin(), c.cons.cols.end(), cols);
    CVM_CATCH(ctx)
}

int cvm_constraints(cvm_context* ctx, const double* x, int flags, double* c_out, double* j_values) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    cvm::ConstraintLayout& L = c.cons;
    if (!L.ready) throw Error(CVM_E_STATE, "constraint layout not built (cvm_constraint_layout)");
    if (!x) throw Error(CVM_E_INVALID, "x is null");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    if (dev && (!c_out || !j_values) && L.nrows) throw Error(CVM_E_INVALID, "device output pointer is null");
    const double* xd = x;
    double* cd = c_out;
    double* jd = j_values;
    cvm::DBuf<double> tc, tj;
    if (!dev) {
        c.d_x.upload(x, 3 * static_cast<size_t>(c.nn), c.stream);
        xd = c.d_x.p;
        tc.alloc(std::max<int64_t>(L.nrows, 1));
        tj.alloc(std::max<int64_t>(L.ntrip, 1));
        cd = tc.p;
        jd = tj.p;
    }
    cvm::launch_constraints(c, xd, cd, jd);
    if (!dev) {
        if (c_out && L.nrows) tc.download(c_out, L.nrows, c.stream);
        if (j_values && L.ntrip) tj.download(j_values, L.ntrip, c.stream);
        CVM_HIP(hipStreamSynchronize(c.stream));
    }
    CVM_CATCH(ctx)
}

int cvm_external_force(cvm_context* ctx, int flags, double* F_ext) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (!F_ext) throw Error(CVM_E_INVALID, "F_ext is null");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    double* Fd = F_ext;
    if (!dev) {
        c.d_F.alloc(std::max<size_t>(3 * static_cast<size_t>(c.nn), 1));
        Fd = c.d_F.p;
    }
    cvm::launch_external_force(c, Fd);
    if (!dev) {
        CVM_HIP(hipMemcpyAsync(F_ext, Fd, 3 * sizeof(double) * c.nn, hipMemcpyDeviceToHost, c.stream));
        CVM_HIP(hipStreamSynchronize(c.stream));
    }
    CVM_CATCH(ctx)
}

int cvm_energies(cvm_context* ctx, const double* x, int flags, double* out3) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (!x || !out3) throw Error(CVM_E_INVALID, "null argument");
    if (!c.have_mat) throw Error(CVM_E_STATE, "laminate not set");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    const double* xd = x;
    if (!dev) {
        c.d_x.upload(x, 3 * static_cast<size_t>(c.nn), c.stream);
        xd = c.d_x.p;
    }
    double* od = out3;
    if (!dev) {
        c.en_out.alloc(3);
        od = c.en_out.p;
    }
    cvm::launch_energies(c, xd, od);
    if (!dev) CVM_HIP(hipMemcpyAsync(out3, od, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));   // the activity copies read host vectors
    CVM_CATCH(ctx)
}

int cvm_fold_angles(cvm_context* ctx, int fold_id, const double* x, int flags, double* theta) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (fold_id < 0 || fold_id >= static_cast<int>(c.folds.size())) throw Error(CVM_E_INVALID, "bad fold id");
    if (!x || !theta) throw Error(CVM_E_INVALID, "null argument");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    const int n = c.doms[c.folds[fold_id].left]->ng;
    const double* xd = x;
    double* td = theta;
    if (!dev) {
        c.d_x.upload(x, 3 * static_cast<size_t>(c.nn), c.stream);
        xd = c.d_x.p;
        c.en_scr.ensure(static_cast<size_t>(std::max(n, 1)));
        td = c.en_scr.p;
    }
    cvm::launch_fold_angles(c, fold_id, xd, td);
    if (!dev && n > 0) CVM_HIP(hipMemcpyAsync(theta, td, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    CVM_CATCH(ctx)
}

int cvm_export_fields(cvm_context* ctx, const double* x, int flags, int nfields, const int* field_codes,
                      double* out) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (!x || !out || nfields < 0 || (nfields > 0 && !field_codes)) throw Error(CVM_E_INVALID, "null argument");
    cvm::Domain* face = nullptr;
    for (cvm::Domain* d : c.doms) if (d->kind == CVM_DOMAIN_FACE) { face = d; break; }
    if (!face) throw Error(CVM_E_STATE, "no face domain");
    const bool dev = (flags & CVM_DEVICE_PTRS) != 0;
    const double* xd = x;
    if (!dev) {
        c.d_x.upload(x, 3 * static_cast<size_t>(c.nn), c.stream);
        xd = c.d_x.p;
    }
    std::vector<int> codes(field_codes, field_codes + nfields);
    if (codes.empty()) codes.push_back(-1);
    c.ex_codes.upload(codes, c.stream);
    const size_t n = static_cast<size_t>(face->ng) * static_cast<size_t>(7 + nfields);
    double* od = out;
    if (!dev) {
        c.ex_out.alloc(std::max<size_t>(n, 1));
        od = c.ex_out.p;
    }
    cvm::launch_export_fields(c, xd, nfields, c.ex_codes.p, od);
    if (!dev && n) CVM_HIP(hipMemcpyAsync(out, od, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    CVM_CATCH(ctx)
}

int cvm_clear_creases(cvm_context* ctx) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (c.batched) throw Error(CVM_E_STATE, "creases of a batched context cannot be replaced");
    c.creases.clear();
    c.creases_dirty = true;
    c.node_sign_creases = -1;
    for (cvm::Domain* d : c.doms) { d->have_sf = false; d->ind_creases = -1; }
    CVM_CATCH(ctx)
}

int cvm_set_tile_shape(cvm_context* ctx, int tile_gps, int union_max, int grid) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (tile_gps < 1 || tile_gps > cvm::kTileGps) throw Error(CVM_E_INVALID, "tile_gps must be in 1..16");
    if (union_max != 0 && (union_max < 8 || union_max > cvm::kTileUnion))
        throw Error(CVM_E_INVALID, "union_max must be 0 (automatic) or in 8..255");
    if (grid < 0) throw Error(CVM_E_INVALID, "grid must be >= 0");
    if (tile_gps != c.tile_gps || union_max != c.union_max) c.plan.ready = false;
    c.tile_gps = tile_gps;
    c.union_max = union_max;
    c.tile_grid = grid;
    CVM_CATCH(ctx)
}

int cvm_get_tile_shape(cvm_context* ctx, int* out4) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (!out4) throw Error(CVM_E_INVALID, "null output");
    out4[0] = c.plan.ready ? c.plan.tile_max_gps : c.tile_gps;
    out4[1] = c.plan.ready ? c.plan.union_max : c.union_max;
    out4[2] = c.tile_grid;
    out4[3] = c.plan.ready ? c.plan.ntiles : 0;
    CVM_CATCH(ctx)
}

int cvm_graph_capture(cvm_context* ctx, int flags, const double* x_dev, double* F_dev, double* K_dev) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    CVM_HIP(hipSetDevice(c.device));
    if (!c.plan.ready) throw Error(CVM_E_STATE, "assembly plan not initialised (cvm_init_assembly)");
    if (!x_dev || !F_dev || ((flags & CVM_TANGENT) && !K_dev)) throw Error(CVM_E_INVALID, "null device pointer");
    if (c.graph_exec) { (void)hipGraphExecDestroy(c.graph_exec); c.graph_exec = nullptr; }
    if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
    // everything the launches need is allocated before the capture starts; an
    // FP32-tangent capture also reserves the FP64 scratch, so that an eager FP64
    // evaluation between replays (a mixed-precision Newton loop's finishing
    // step) does not reallocate the captured scratch
    const int eflags = flags & ~CVM_DEVICE_PTRS;
    cvm::reserve_workspace(c, c.ws, eflags & ~CVM_FP32_TANGENT);
    cvm::reserve_workspace(c, c.ws, eflags);
    CVM_HIP(hipStreamSynchronize(c.stream));
    const bool timed = c.ws.timed;
    c.ws.timed = false;
    cvm::LaunchSig sig;
    CVM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
        cvm::launch_assembly(c, c.ws, eflags, x_dev, F_dev, K_dev, &sig);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c.ws.timed = timed;
        throw;
    }
    c.ws.timed = timed;
    CVM_HIP(hipStreamEndCapture(c.stream, &c.graph));
    CVM_HIP(hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0));
    c.graph_sig = sig.h;
    c.graph_flags = eflags;
    c.graph_x = x_dev; c.graph_F = F_dev; c.graph_K = K_dev;
    CVM_CATCH(ctx)
}

int cvm_graph_launch(cvm_context* ctx) {
    CVM_TRY(ctx)
    Context& c = ctx->c;
    if (!c.graph_exec) throw Error(CVM_E_STATE, "no captured graph (cvm_graph_capture)");
    if (!c.plan.ready) throw Error(CVM_E_STATE, "the assembly plan changed since the capture");
    // the graph is valid while the eager path would issue the captured launches
    cvm::LaunchSig sig;
    sig.dry = true;
    cvm::launch_assembly(c, c.ws, c.graph_flags, c.graph_x, c.graph_F, c.graph_K, &sig);
    if (sig.grow || sig.h != c.graph_sig)
        throw Error(CVM_E_STATE, "the assembly changed since the capture (plan, laminate, shape functions, "
                                 "workspace or tile grid): capture again (cvm_graph_capture)");
    CVM_HIP(hipSetDevice(c.device));
    const int slot = cvm::timing_begin(c, "assembly_graph");
    CVM_HIP(hipGraphLaunch(c.graph_exec, c.stream));
    cvm::timing_end(c, slot);
    CVM_CATCH(ctx)
}

int cvm_synchronize(cvm_context* ctx) {
    CVM_TRY(ctx)
    CVM_HIP(hipStreamSynchronize(ctx->c.stream));
    CVM_CATCH(ctx)
}

int cvm_set_profiling(cvm_context* ctx, int on) {
    CVM_TRY(ctx)
    cvm::timing_collect(ctx->c);
    ctx->c.timing.on = on != 0;
    CVM_CATCH(ctx)
}

int cvm_kernel_count(cvm_context* ctx) {
    if (!ctx) return 0;
    try { cvm::timing_collect(ctx->c); } catch (...) {}
    return static_cast<int>(ctx->c.timing.names.size());
}

int cvm_kernel_stats(cvm_context* ctx, int i, char* name, int name_cap, double* total_ms, int64_t* launches) {
    CVM_TRY(ctx)
    cvm::Timing& t = ctx->c.timing;
    cvm::timing_collect(ctx->c);
    if (i < 0 || i >= static_cast<int>(t.names.size())) throw Error(CVM_E_INVALID, "bad kernel index");
    if (name && name_cap > 0) std::snprintf(name, name_cap, "%s", t.names[i].c_str());
    if (total_ms) *total_ms = t.total_ms[i];
    if (launches) *launches = t.count[i];
    CVM_CATCH(ctx)
}

int cvm_reset_profiling(cvm_context* ctx) {
    CVM_TRY(ctx)
    cvm::timing_collect(ctx->c);
    for (auto& v : ctx->c.timing.total_ms) v = 0.0;
    for (auto& v : ctx->c.timing.count) v = 0;
    CVM_CATCH(ctx)
}

}  // extern "C"
