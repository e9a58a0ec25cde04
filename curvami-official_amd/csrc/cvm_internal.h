// Internal definitions of the curvami GPU library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/curvami_gpu.h"

namespace cvm {

// Error carrying a C-ABI status code; converted to a status at the ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CVM_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::cvm::Error(CVM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Owning device allocation.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(CVM_E_NO_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    // grow-only: keeps a larger buffer (workspaces shared by sheets of different sizes)
    void ensure(size_t count) {
        if (count <= n && (p || count == 0)) return;
        alloc(count);
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) CVM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) CVM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct Crease {
    std::vector<double> vx, vy;
    int ext_s = 0, ext_e = 0, closed = 0;
};

// Shape-function cache of one domain, device resident, pair order = FRNN order
// (GP-major, neighbours in the search order of GlobalFRNN::execute).
// BoundaryCondition of a domain (plate/PlateData.h:36-50): active bit 0 u,
// 1 v, 2 w, 3 normal rotation; dead force per GP; edge moment per unit length.
struct BoundaryCond {
    int active = 0;
    double val[4] = {0, 0, 0, 0};
    double force[3] = {0, 0, 0};
    double moment = 0.0;
};

// A sheet of a batched context: nodes [node0, node0 + nn), creases
// [crease0, crease0 + ncreases), its own support radius. A context set up with
// cvm_set_nodes has one implicit sheet that covers everything.
struct SheetSeg {
    int node0 = 0, nn = 0, crease0 = 0, ncreases = 0;
    double rho = 0.0;
};

// Device view of a sheet for one domain's neighbour search and shape functions:
// the sheet's cell grid (keys koff + y * cols + x), radius and crease range.
// GPs per block of the shape-function kernel (one wave)
constexpr int RKPM_BLOCK = 64;

struct SheetDev {
    double minx = 0.0, miny = 0.0, cell = 0.0, rho = 0.0, rho2 = 0.0;
    double irho = 0.0, irho2 = 0.0;     // 1 / rho, 1 / rho^2 (IEEE, host)
    long long koff = 0;
    int cols = 0, nkeys = 0, nc = 0, crease0 = 0;
};

struct Domain {
    int kind = CVM_DOMAIN_FACE;
    int ng = 0;
    std::vector<int> h_gsheet;     // per GP: sheet (batched contexts only; empty = sheet 0)
    DBuf<int> d_gsheet;
    bool dirty = false;            // host GP arrays not yet uploaded (batched appends)
    DBuf<unsigned char> d_sheets;  // SheetDev per sheet, built by the neighbour search
    std::vector<SheetDev> h_sheets;  // host copy; crease ranges refreshed before each shape-function launch
    std::vector<double> hx, hy, hw, hnx, hny, htx, hty;   // host copies of the GP data
    DBuf<double> gx, gy, d_w, d_tx, d_ty, d_nx, d_ny;
    BoundaryCond bc;
    uint64_t gen = 0;         // bumped whenever neighbours or shape functions change
    // neighbour lists
    double rho = 0.0;
    int64_t npairs = -1;
    DBuf<int> off;            // ng + 1
    DBuf<int> ids;            // npairs
    DBuf<double> s;           // npairs
    std::vector<int> h_off, h_ids;     // host mirrors (needed for the pattern plan)
    // the node cell grid of the last search (FRNN.h:30-51), kept for
    // cvm_refresh_neighbours; valid for grid_nn nodes and grid_ng GPs
    DBuf<int> cell_start, cell_nodes;
    DBuf<int> cnt;            // per GP neighbour counts of the search
    DBuf<unsigned char> scan_tmp;
    int grid_nn = -1, grid_ng = -1;
    // The same lists interleaved per block of RKPM_BLOCK GPs for the moment
    // pass, where each lane walks its own GP's pairs: the t-th pair of GP lane
    // l of block b sits at (il_row[b] + t) * RKPM_BLOCK + l, so one wave-wide
    // load reads consecutive words. Padding slots (t >= count) are never read.
    DBuf<int> il_row;         // nblocks + 1 (rows of RKPM_BLOCK slots)
    DBuf<int> il_ids;         // il_row[nblocks] * RKPM_BLOCK
    DBuf<double> il_s;
    DBuf<uint8_t> pair_lgp;   // npairs: GP of the pair minus the first GP of its RKPM block
    // shape functions: psi, dpx, dpy, dxx, dyy, dxy (SoA, npairs each)
    DBuf<double> sf;          // 6 * npairs
    int order = 0, kernel = 0;
    bool have_sf = false;
    // per-GP crease indicator data: 6 doubles per (crease, GP); valid for `ind_creases` creases
    DBuf<double> ind;
    int ind_creases = -1;
};

// Natural curvature band (Kinematics::NaturalCurvature, plate/Kinematics.h:9-45):
// kappa_bar(0, 0) = sign * theta0 * exp(-((X - x0) / width)^2) / (sqrt(pi) width)
// at a face GP's reference position, added to E1(0, 0) in the resultants and
// the bending energy. Inactive by default.
struct NatCurv {
    bool active = false;
    double x0 = 0.0, width = 1.0, theta0 = 0.0, sign = -1.0;
    double kappa00(double X) const {
        if (!active) return 0.0;
        const double t = (X - x0) / width;
        return sign * theta0 * (std::exp(-t * t) / (1.7724538509055160273 * width));
    }
};

struct Fold {
    int left = -1, right = -1;
    double Kt = 0.0, Ang0 = 0.0;
    bool rigid = false;           // per-fold rigid hinge: constraint rows instead of the elastic hinge
    std::vector<double> theta0;   // per crease GP (empty: Ang0)
    std::vector<char> active;     // per crease GP (empty: all active)
    DBuf<double> d_theta0;
};

// Constraint rows of Meshless::evaluateConstraints (plate/Meshless.h:129-195):
// one job per Dirichlet GP or active crease GP, with its first row and first
// triplet; the triplet pattern (rows, cols) is fixed by the neighbour lists.
struct ConstraintLayout {
    bool ready = false;
    bool rigid = false;
    int64_t nrows = 0, ntrip = 0;
    std::vector<int> rows, cols;
    std::vector<uint64_t> gens;   // domain generations the layout was built from
    int njobs = 0;
    DBuf<int> d_job;              // 6 ints per job: kind, a, b, gp, row0, trip0
    DBuf<unsigned char> d_dom;    // per domain: device descriptor
    DBuf<unsigned char> d_fold;   // per fold: device descriptor
    // external force: contributions grouped by dof, in the reference's order
    bool fext_ready = false;
    std::vector<uint64_t> fext_gens;
    DBuf<int> d_fx_off;           // 3 nn + 1
    DBuf<int> d_fx_rec;           // 4 ints per contribution: dom, pair, gp, type
    DBuf<double> d_fx_val;        // 3 doubles per contribution: coefficient, n_out x, n_out y
};

// Signature of the launches one assembly issues: each kernel, its grid, block
// and LDS size and every argument by value. A captured graph replays exactly
// the launches of its capture, so it stays valid only while an eager call with
// the same flags and pointers would issue the same signature (same plan
// buffers and tile counts, laminate, workspace buffers and tile grid).
struct LaunchSig {
    bool dry = false;    // record the signature only: no launch, no allocation
    bool grow = false;   // (dry) a workspace buffer would have to be reallocated
    uint64_t h = 1469598103934665603ull;
    void mix(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    }
};

// What one assembly in flight needs besides the plan: its stream, the tile
// scratch rows, the per-union-node residual rows and the hinge records. The
// context owns one (on its own stream); a serving loop owns one per slot.
struct Workspace {
    hipStream_t stream = nullptr;
    bool timed = false;          // kernel timings go to the context's Timing
    DBuf<double> scr;            // 9 * nscratch values of the tangent type
    DBuf<double> fscr;           // 3 * nfscratch
    DBuf<double> fold_gp;        // per crease GP record
    DBuf<double> fold_nd;        // per (crease GP, node) record
};

// Assembly plan: tiles of face GPs, block pattern, contributor lists.
struct Plan {
    bool ready = false;
    int nn = 0;
    // block-upper pattern at node level: CSC of blocks (union of face and fold pairs)
    std::vector<int> col_ptr;       // nn + 1 (blocks per column node)
    std::vector<int> col_rows;      // nblocks: row node of each block (ascending in a column)
    std::vector<char> blk_face;     // 1 if block is in the face pattern
    int64_t nblocks = 0;
    // face-only pattern (CSC of scalars) and union pattern
    std::vector<int> face_outer, face_inner, uni_outer, uni_inner;
    // per block: base index and column stride in both patterns (-1: absent)
    DBuf<int64_t> d_blk_base_u, d_blk_base_f;
    DBuf<int> d_blk_stride_u, d_blk_stride_f;
    // tiles
    int ntiles = 0;
    int tile_max_gps = 16, union_max = 64;
    int max_nbr = 0;            // largest face-GP neighbour count
    DBuf<int> d_tile_gp;        // ntiles + 1
    DBuf<int> d_tile_uoff;      // ntiles + 1
    DBuf<int> d_tile_unodes;    // union node ids
    DBuf<int64_t> d_tile_toff;  // ntiles + 1 (touched pairs)
    DBuf<uint16_t> d_tile_tpairs;   // (ua << 8) | ub; per tile its paired slots first
    DBuf<int> d_tile_nd;        // ntiles: paired slots of the tile (pairs of slots)
    DBuf<uint8_t> d_pair_uloc;  // per face pair: union-local index
    DBuf<uint8_t> d_pair_urank; // per face pair: rank of its node among the GP's neighbours in union order
    int64_t nscratch = 0;       // total touched pairs
    int64_t nfscratch = 0;      // total union entries
    // K contributors: per block, a contiguous segment of scratch rows (ascending
    // tile order); slot_dst maps a tile's touched slot to its row
    DBuf<int64_t> d_kc_off;
    DBuf<uint32_t> d_slot_dst;
    // F contributors: per node, fscratch indices
    DBuf<int> d_fc_off;
    DBuf<int> d_fc_idx;
    // folds: crease GPs aggregated per node
    int nfgp = 0;                       // total crease GPs over folds
    DBuf<int> d_fgp_fold;               // fold of each crease GP
    DBuf<int> d_fgp_local;              // index within its fold
    DBuf<int> d_fgp_noff;               // per crease GP: offsets into node set (nfgp + 1)
    DBuf<int> d_fgp_nodes;              // node ids of L u R (sorted unique)
    DBuf<int> d_fgp_lpos;               // per L neighbour entry: position in node set (by crease GP)
    DBuf<int> d_fgp_rpos;               // per R neighbour entry
    DBuf<int> d_fgp_lpo, d_fgp_rpo;     // per crease GP: offsets into lpos / rpos
    DBuf<int> d_fnd_gp;                 // node-set entry -> crease GP
    int64_t nfnodes = 0;                // node-set entries over all crease GPs
    DBuf<unsigned char> d_folddev;      // per fold: device descriptor (shape functions, weights, theta0)
    DBuf<int64_t> d_kf_off;             // per block: fold contributor offsets
    DBuf<int> d_kf_idx;                 // per contribution: node-set entries of lo and hi
    DBuf<int> d_ff_off;                 // per node: fold contributions (nodeset entry indices)
    DBuf<int> d_ff_idx;
    int max_touched = 0, max_union = 0, rec_cap = 0;
    int face_ng = 0;
    int64_t face_npairs = 0;
    std::vector<std::vector<int>> face_rows;   // face pattern rows per column (csrMapping)
};

struct Timing {
    bool on = false;
    std::vector<std::string> names;
    std::vector<hipEvent_t> ev0, ev1;
    std::vector<int> pending;   // kernel id per event pair recorded since last read
    std::vector<double> total_ms;
    std::vector<int64_t> count;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // nodes
    int nn = 0;
    std::vector<double> hX, hY;
    DBuf<double> X, Y;
    // material
    double A[9] = {0}, B[9] = {0}, D[9] = {0};
    bool have_mat = false;
    // natural curvature and its per-face-GP kappa_bar(0, 0) on the device
    NatCurv nat;
    DBuf<double> d_kbar;
    int kbar_ng = -1;             // face GP count d_kbar was formed for (-1: stale)
    // creases (host) + device packed vertices
    std::vector<Crease> creases;
    DBuf<double> d_cv;          // packed vertices x,y
    DBuf<int> d_cinfo;          // per crease: offset, count, ext_s, ext_e, closed, first group, groups
    DBuf<double> d_cgrp;        // per group of 16 segments: bounding box and first vertex
    DBuf<double> d_cmax;        // per crease: largest |vertex coordinate|
    DBuf<int8_t> d_node_sign;   // per (crease, node): +1 / -1
    std::vector<int> h_crange;  // per sheet: first crease, crease count (node-sign launch)
    DBuf<int> d_crange;
    DBuf<int> rf_flags;         // neighbour refresh: per domain, lists changed
    DBuf<unsigned char> rf_tab, ind_tab;      // refresh and indicator domain tables
    std::vector<unsigned char> rf_tab_host, ind_tab_host;
    bool creases_dirty = true;
    int node_sign_creases = -1;   // crease count the node signs were computed for (-1: stale)
    // sheets (batched contexts): nodes and creases per sheet, per-node sheet
    std::vector<SheetSeg> sheets;
    bool batched = false;
    std::vector<int> h_nsheet;
    DBuf<int> d_nsheet;
    bool nodes_dirty = false;      // host node arrays not yet uploaded (batched appends)
    int max_sheet_creases() const {
        if (!batched) return static_cast<int>(creases.size());
        int m = 0;
        for (const SheetSeg& s : sheets) m = std::max(m, s.ncreases);
        return m;
    }
    std::vector<Domain*> doms;
    std::vector<Fold> folds;
    Plan plan;
    ConstraintLayout cons;
    DBuf<int> d_status;         // RKPM status words, 4 per domain
    DBuf<double> rk_cj;         // RKPM per-GP coefficient records (grow-only, reused launch by launch)
    DBuf<unsigned char> rk_tab[2];            // RKPM domain tables: all domains, single domain
    std::vector<unsigned char> rk_tab_host[2];
    // assembly buffers
    DBuf<double> d_x;           // 3 nn displacements
    DBuf<double> en_scr;        // per-GP energies (face membrane, face bending, crease hinge)
    DBuf<char> en_active;       // crease-GP activity copies for the hinge energy
    DBuf<double> en_out;        // 3 energies
    DBuf<int> ex_codes;         // export field codes
    DBuf<double> ex_out;        // export columns (ng x (7 + fields))
    DBuf<double> d_F;           // 3 nn
    DBuf<double> d_K;           // nnz (union)
    Workspace ws;               // the context's own assembly workspace (its stream)
    // A second stream (default priority) for work that runs beside the
    // context's stream: the crease indicators and node sides of a neighbour
    // refresh, beside its verification pass. fork / join order the two.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    size_t lds_limit = 65536;
    Timing timing;
    int tile_gps = 16;          // GPs per assembly tile (at most kTileGps)
    int union_max = 0;          // union nodes per tile (at most kTileUnion); 0: chosen from the neighbour counts
    int tile_grid = 0;          // tile kernel grid: 0 = one workgroup per tile, else persistent
    // captured assembly (cvm_graph_capture)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_sig = 0;     // launch signature at capture
    int graph_flags = 0;
    const double* graph_x = nullptr;
    double* graph_F = nullptr;
    double* graph_K = nullptr;
    ~Context();
};

// kernel launchers (implemented in the .hip files)
// Indicator values and node sides (when stale). Uploads go on the context's
// stream; the kernel runs on kst (default: the context's stream) after the
// uploads, and the caller orders later work of the context's stream after it.
void launch_indicators(Context& c, const std::vector<Domain*>& ds, hipStream_t kst = nullptr);
void launch_node_signs(Context& c, hipStream_t kst = nullptr);
// The context's side stream and its fork / join events (created on first use).
hipStream_t side_stream(Context& c);
void launch_frnn(Context& c, Domain& d, double rho);
// Repeats the search of every searched domain on the device (same nodes, GPs
// and radius): the lists are rewritten in place; per domain, whether its lists
// changed (lists, mirrors and interleaved copies then replaced).
// beside (optional) enqueues more work on the side stream after the crease
// indicators, beside the verification pass; it is called only when the
// verification runs (some domain keeps its grid), and its work is complete
// when the call returns.
// joined (optional) enqueues work on the context's stream after the side
// stream's work, before the one synchronisation of the call (read-backs).
std::vector<char> refresh_frnn_all(Context& c, const std::function<void(hipStream_t)>& beside = nullptr,
                                   const std::function<void()>& joined = nullptr);
void launch_rkpm(Context& c, Domain& d, double rho, int order, int kernel, int* code, int* nbad, int* first);
void launch_rkpm_all(Context& c, int order, int kernel, int* dom, int* code, int* nbad, int* first);
// The shape functions of every domain enqueued with their kernels on stream
// kst (default: the context's stream), statuses in the context's status
// words; then, after they completed, the statuses read back and the first
// failing domain (in domain order) and its status.
struct RkpmBatch {
    std::vector<size_t> order_of;   // domain of each launch slot
    int order = 0, kernel = 0;
};
RkpmBatch enqueue_rkpm_all(Context& c, int order, int kernel, hipStream_t kst = nullptr);
void finish_rkpm_all(Context& c, const RkpmBatch& b, int* dom, int* code, int* nbad, int* first);
// The status words of a batch into host memory hs (4 per domain) on the
// context's stream (asynchronous), and the statuses once read.
void read_rkpm_status(Context& c, int* hs);
void finish_rkpm_all(Context& c, const RkpmBatch& b, const int* hs, int* dom, int* code, int* nbad, int* first);
// Orders stream kst after the work enqueued so far on the context's stream;
// returns the stream to launch on (the context's stream when kst is null).
hipStream_t fork_to(Context& c, hipStream_t kst);
// Workspace buffers given by address (the allocation-free entry point's static
// scratch) instead of a Workspace's own buffers.
struct WsPtrs {
    double* scr = nullptr;
    double* fscr = nullptr;
    double* fold_gp = nullptr;
    double* fold_nd = nullptr;
};
void launch_assembly(Context& c, Workspace& ws, int flags, const double* xdev, double* Fdev, double* Kdev,
                     LaunchSig* sig = nullptr, const WsPtrs* ext = nullptr, bool tile_done = false);
// Tiles of several evaluations (contexts) in one launch, k_tile_multi: the
// serving loop's requests admitted together. Evaluations share a launch when
// their keys are equal (tangent type, laminate class, neighbour width; -1: a
// persistent tile grid, never shared); a launch takes at most kTileJobsMax
// of them (the job table is a kernel argument). The evaluations' remaining
// stages are launch_assembly(..., tile_done = true) on the same stream.
constexpr int kTileJobsMax = 8;
struct TileGroupItem {
    Context* c;
    int flags;
    const double* x;
    double* scr;
    double* fscr;
};
int tile_group_key(const Context& c, int flags);
void launch_tile_group(const TileGroupItem* items, int n, hipStream_t st);
// The allocation-free evaluation (cvm_assemble_noalloc): the workspace lies in
// the library's static device scratch of the context's GPU.
void noalloc_scratch_init(int device);
void launch_assembly_noalloc(Context& c, int flags, const double* xdev, double* Fdev, double* Kdev);
size_t assembly_workspace_bytes(const Context& c, int flags);
void reserve_workspace(const Context& c, Workspace& ws, int flags);
// true if reserve_workspace(c, ws, flags) would reallocate a buffer
bool workspace_would_grow(const Context& c, const Workspace& ws, int flags);
// element counts (doubles) of the workspace buffers fscr, scr (0 without the
// tangent), fold_gp, fold_nd for an evaluation with these flags
void workspace_counts(const Plan& P, int flags, size_t (&n)[4]);
void upload_fold_descriptors(Context& c);
// Device pointer of kappa_bar(0, 0) per face GP, or null with the natural
// curvature inactive; forms and uploads it when stale (allocates) unless
// allow_alloc is false, which then refuses with CVM_E_STATE.
const double* natural_kappa(Context& c, bool allow_alloc);
size_t tile_lds_bytes(int rec_cap, int max_union, bool fp32, bool wide);
int tile_umax(int max_union);
constexpr int kTileGps = 16;    // GPs per assembly tile
constexpr int kTileUnion = 255; // union nodes per assembly tile (uint8 union indices)

// timing helpers
// per-kernel event timing on stream st (default: the context's stream); a
// kernel timed on another stream is joined into the context's stream before
// the timings are read
int timing_begin(Context& c, const char* name, hipStream_t st = nullptr);
void timing_end(Context& c, int slot, hipStream_t st = nullptr);

void build_plan(Context& c);
void sync_nodes(Context& c);
void sync_domain(Context& c, Domain& d);
void build_constraint_layout(Context& c, bool rigid);
void launch_constraints(Context& c, const double* xdev, double* cdev, double* jdev);
void launch_external_force(Context& c, double* Fdev);
void launch_energies(Context& c, const double* xdev, double* out3_dev);
void launch_fold_angles(Context& c, int fold_id, const double* xdev, double* theta_dev);
void launch_export_fields(Context& c, const double* xdev, int nf, const int* fcode_dev, double* out_dev);

}  // namespace cvm

// The opaque handle of the C ABI.
struct cvm_context {
    cvm::Context c;
};

