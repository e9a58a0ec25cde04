// Serving loop: residual / tangent evaluation requests on device-resident
// sheets, evaluated concurrently on a pool of HIP streams.
//
//  * Sheets are prepared contexts (nodes, neighbour lists, shape functions and
//    the assembly plan stay in HBM); a request names a sheet and a displacement.
//  * Slots: up to max_inflight evaluations are in flight at once, each on its own
//    stream with its own workspace (scratch rows, x, F, K, pinned staging).
//    Workspaces are provisioned while the server is idle (at the first request
//    after sheets were registered): slot by slot, each for the largest requests
//    that still fit the HBM budget left (the first slots for every registered
//    sheet, later ones for the smaller sheets); the dispatcher never allocates
//    or frees device memory (hipFree would synchronise the device under
//    requests in flight).
//    Requests that arrive while others run are admitted as soon as a slot that
//    holds them is free, so their kernels overlap on the device (continuous
//    batching).
//  * Requests admitted in the same dispatcher pass (several slots free at
//    once, as when a group of requests completes under load) share their tile
//    launch: the Gauss-point tiles of all their sheets run as one k_tile_multi
//    launch over a job table (a kernel argument; at most 8 requests per
//    launch), then each request's hinge, F and K reductions follow on the
//    same stream. Requests of different tangent types, laminate classes or
//    neighbour widths go into separate groups.
//  * A request's arithmetic does not depend on what else is in flight: each
//    evaluation reads only its sheet's plan and writes only its own workspace,
//    and a tile computes the same in a shared launch as in its own, so results
//    are bit-identical alone and in any mix.
//  * One dispatcher thread admits, launches and retires requests (event polling);
//    submit and wait may be called from any host thread.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "cvm_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct SheetRef {
    cvm_context* ctx = nullptr;
    int nn = 0;
    int64_t nnz_union = 0, nnz_face = 0;
};

struct Request {
    int sheet = 0;
    int flags = 0;
    std::vector<double> x;          // copied at submit
    double* F_host = nullptr;
    double* K_dev = nullptr;        // caller's device buffer, or null (slot buffer)
    size_t bytes = 0;               // workspace bytes charged against the budget
    int slot = -1;
    int status = CVM_OK;
    std::string err;
    bool done = false;
    Clock::time_point t_submit, t_start, t_done;
};

struct Slot {
    cvm::Workspace ws;
    cvm::DBuf<double> x, F, K;
    double* hx = nullptr;           // pinned staging
    double* hF = nullptr;
    size_t hcap = 0;
    hipEvent_t ev = nullptr;
    int64_t req = -1;

    size_t bytes() const {
        return sizeof(double) * (ws.scr.n + ws.fscr.n + ws.fold_gp.n + ws.fold_nd.n + x.n + F.n + K.n);
    }
};

}  // namespace

struct cvm_server {
    int device = 0;
    int64_t budget = 0;
    std::vector<SheetRef> sheets;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_id = 1;
    bool batching = true;            // requests admitted together share their tile launch
    int64_t groups = 0, grouped = 0; // shared tile launches and the requests in them
    int64_t in_use = 0;              // bytes charged by in-flight requests
    int64_t completed = 0;
    int inflight = 0;
    bool stop = false;
    std::string err;
    std::thread worker;

    size_t request_bytes(const SheetRef& s, int flags, bool own_K) const {
        size_t b = cvm::assembly_workspace_bytes(s.ctx->c, flags);
        b += 2 * sizeof(double) * 3 * static_cast<size_t>(s.nn);          // x, F
        if ((flags & CVM_TANGENT) && own_K)
            b += sizeof(double) * static_cast<size_t>((flags & CVM_ELASTIC_FOLDS) ? s.nnz_union : s.nnz_face);
        return b;
    }

    // allocated workspace bytes over all slots
    size_t allocated() const {
        size_t t = 0;
        for (const Slot& s : slots) t += s.bytes();
        return t;
    }

    // flags of a sheet's largest request: FP64 tangent with the hinge, K in the slot
    static constexpr int kFullFlags = CVM_RESIDUAL | CVM_TANGENT | CVM_ELASTIC_FOLDS;

    // true if slot sl holds a request of sheet sh with these flags without growing
    static bool fits(const Slot& sl, const SheetRef& sh, int flags, bool own_K) {
        const size_t n3 = 3 * static_cast<size_t>(sh.nn);
        const int64_t nnz = (flags & CVM_ELASTIC_FOLDS) ? sh.nnz_union : sh.nnz_face;
        if (cvm::workspace_would_grow(sh.ctx->c, sl.ws, flags)) return false;
        if (sl.x.n < n3 || sl.F.n < n3 || sl.hcap < n3) return false;
        if ((flags & CVM_TANGENT) && own_K && sl.K.n < static_cast<size_t>(std::max<int64_t>(nnz, 1))) return false;
        return true;
    }

    // element counts of a slot's buffers: fscr, scr, fold_gp, fold_nd, x (and F), K
    struct SlotSize {
        size_t n[6] = {0, 0, 0, 0, 0, 0};
        size_t bytes() const { return sizeof(double) * (n[0] + n[1] + n[2] + n[3] + 2 * n[4] + n[5]); }
        void cover(const SheetRef& sh) {
            size_t w[4];
            cvm::workspace_counts(sh.ctx->c.plan, kFullFlags, w);
            for (int k = 0; k < 4; ++k) n[k] = std::max(n[k], w[k]);
            n[4] = std::max(n[4], 3 * static_cast<size_t>(sh.nn));
            n[5] = std::max(n[5], static_cast<size_t>(std::max<int64_t>(sh.nnz_union, 1)));
        }
    };

    bool planned = true;             // false after a sheet is registered, until the slots are provisioned

    // Provisions the slots (server idle): the sheets by their largest request,
    // descending; slot after slot takes the buffers that hold every sheet from
    // the largest one that still fits the budget left down to the smallest.
    void plan_slots() {
        (void)hipSetDevice(device);
        for (Slot& sl : slots) {
            sl.ws.scr.release(); sl.ws.fscr.release(); sl.ws.fold_gp.release(); sl.ws.fold_nd.release();
            sl.x.release(); sl.F.release(); sl.K.release();
        }
        std::vector<int> order(sheets.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = static_cast<int>(k);
        std::vector<size_t> need(sheets.size());
        for (size_t k = 0; k < sheets.size(); ++k) need[k] = request_bytes(sheets[k], kFullFlags, true);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return need[a] > need[b]; });
        // suffix covers: buffers holding the sheets order[i..]
        std::vector<SlotSize> cover(order.size() + 1);
        for (size_t i = order.size(); i-- > 0;) {
            cover[i] = cover[i + 1];
            cover[i].cover(sheets[order[i]]);
        }
        size_t left = static_cast<size_t>(budget);
        for (Slot& sl : slots) {
            size_t i = 0;
            while (i < order.size() && cover[i].bytes() > left) ++i;
            if (i == order.size()) break;
            const SlotSize& z = cover[i];
            sl.ws.fscr.alloc(z.n[0]);
            sl.ws.scr.alloc(z.n[1]);
            sl.ws.fold_gp.alloc(z.n[2]);
            sl.ws.fold_nd.alloc(z.n[3]);
            sl.x.alloc(z.n[4]);
            sl.F.alloc(z.n[4]);
            sl.K.alloc(z.n[5]);
            if (sl.hcap < z.n[4]) {
                if (sl.hx) (void)hipHostFree(sl.hx);
                if (sl.hF) (void)hipHostFree(sl.hF);
                sl.hx = sl.hF = nullptr;
                sl.hcap = 0;
                CVM_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hx), z.n[4] * sizeof(double), hipHostMallocDefault));
                CVM_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hF), z.n[4] * sizeof(double), hipHostMallocDefault));
                sl.hcap = z.n[4];
            }
            left -= z.bytes();
        }
        planned = true;
    }

    void launch(int64_t id, Request& r, int si) {
        Slot& sl = slots[si];
        const SheetRef& sh = sheets[r.sheet];
        cvm::Context& c = sh.ctx->c;
        const size_t n3 = 3 * static_cast<size_t>(sh.nn);
        const bool own_K = r.K_dev == nullptr;
        if (!fits(sl, sh, r.flags, own_K)) throw cvm::Error(CVM_E_STATE, "serving slot not provisioned for this request");
        std::memcpy(sl.hx, r.x.data(), n3 * sizeof(double));
        CVM_HIP(hipMemcpyAsync(sl.x.p, sl.hx, n3 * sizeof(double), hipMemcpyHostToDevice, sl.ws.stream));
        double* Kd = (r.flags & CVM_TANGENT) ? (own_K ? sl.K.p : r.K_dev) : nullptr;
        cvm::launch_assembly(c, sl.ws, r.flags & ~CVM_DEVICE_PTRS, sl.x.p, sl.F.p, Kd);
        CVM_HIP(hipMemcpyAsync(sl.hF, sl.F.p, n3 * sizeof(double), hipMemcpyDeviceToHost, sl.ws.stream));
        CVM_HIP(hipEventRecord(sl.ev, sl.ws.stream));
        sl.req = id;
        r.slot = si;
        r.t_start = Clock::now();
    }

    // Requests admitted together (slots reserved): groups of equal tile keys of two
    // or more (at most kTileJobsMax per launch) share one tile launch on the first
    // request's stream; the rest run alone on their own slots' streams.
    void launch_batch(const std::vector<std::pair<int64_t, int>>& batch) {
        std::vector<int> key(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) {
            const Request& r = reqs[batch[i].first];
            key[i] = batching ? cvm::tile_group_key(sheets[r.sheet].ctx->c, r.flags & ~CVM_DEVICE_PTRS) : -1;
        }
        std::vector<char> done(batch.size(), 0);
        for (size_t i = 0; i < batch.size(); ++i) {
            if (done[i]) continue;
            std::vector<size_t> grp{i};
            if (key[i] >= 0)
                for (size_t k = i + 1; k < batch.size() && grp.size() < static_cast<size_t>(cvm::kTileJobsMax); ++k)
                    if (!done[k] && key[k] == key[i]) grp.push_back(k);
            for (size_t k : grp) done[k] = 1;
            if (grp.size() == 1) {
                launch_one(batch[i].first, batch[i].second);
            } else {
                std::vector<std::pair<int64_t, int>> g;
                for (size_t k : grp) g.push_back(batch[k]);
                launch_group(g);
            }
        }
    }

    void launch_one(int64_t id, int si) {
        Request& r = reqs[id];
        try {
            launch(id, r, si);
        } catch (const cvm::Error& e) {
            unreserve(id, si, e);
        } catch (const std::exception& e) {
            unreserve(id, si, cvm::Error(CVM_E_NO_MEMORY, e.what()));
        }
    }

    void launch_group(const std::vector<std::pair<int64_t, int>>& g) {
        Slot& lead = slots[g[0].second];
        const hipStream_t st = lead.ws.stream;
        try {
            std::vector<cvm::TileGroupItem> items(g.size());
            for (size_t k = 0; k < g.size(); ++k) {
                Request& r = reqs[g[k].first];
                Slot& sl = slots[g[k].second];
                const SheetRef& sh = sheets[r.sheet];
                if (!fits(sl, sh, r.flags, r.K_dev == nullptr))
                    throw cvm::Error(CVM_E_STATE, "serving slot not provisioned for this request");
                const size_t n3 = 3 * static_cast<size_t>(sh.nn);
                std::memcpy(sl.hx, r.x.data(), n3 * sizeof(double));
                CVM_HIP(hipMemcpyAsync(sl.x.p, sl.hx, n3 * sizeof(double), hipMemcpyHostToDevice, st));
                items[k] = cvm::TileGroupItem{&sh.ctx->c, r.flags & ~CVM_DEVICE_PTRS, sl.x.p, sl.ws.scr.p, sl.ws.fscr.p};
            }
            cvm::launch_tile_group(items.data(), static_cast<int>(items.size()), st);
            for (size_t k = 0; k < g.size(); ++k) {
                Request& r = reqs[g[k].first];
                Slot& sl = slots[g[k].second];
                const SheetRef& sh = sheets[r.sheet];
                const size_t n3 = 3 * static_cast<size_t>(sh.nn);
                const cvm::WsPtrs wp{sl.ws.scr.p, sl.ws.fscr.p, sl.ws.fold_gp.p, sl.ws.fold_nd.p};
                double* Kd = (r.flags & CVM_TANGENT) ? (r.K_dev ? r.K_dev : sl.K.p) : nullptr;
                cvm::launch_assembly(sh.ctx->c, lead.ws, r.flags & ~CVM_DEVICE_PTRS, sl.x.p, sl.F.p, Kd, nullptr, &wp, true);
                CVM_HIP(hipMemcpyAsync(sl.hF, sl.F.p, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
                CVM_HIP(hipEventRecord(sl.ev, st));
                r.t_start = Clock::now();
            }
            ++groups;
            grouped += static_cast<int64_t>(g.size());
        } catch (const std::exception& e) {
            // nothing of the group is reported: wait for what was queued, then fail all of it
            (void)hipStreamSynchronize(st);
            const cvm::Error* ce = dynamic_cast<const cvm::Error*>(&e);
            for (const auto& q : g) unreserve(q.first, q.second, ce ? *ce : cvm::Error(CVM_E_NO_MEMORY, e.what()));
        }
    }

    // a reserved request that could not be launched: fail it and free its slot
    void unreserve(int64_t id, int si, const cvm::Error& e) {
        Request& r = reqs[id];
        slots[si].req = -1;
        r.slot = -1;
        in_use -= static_cast<int64_t>(r.bytes);
        --inflight;
        fail(id, r, e);
        cv_done.notify_all();
    }

    void retire(int si) {
        Slot& sl = slots[si];
        Request& r = reqs[sl.req];
        const SheetRef& sh = sheets[r.sheet];
        if (r.F_host) std::memcpy(r.F_host, sl.hF, 3 * sizeof(double) * sh.nn);
        r.t_done = Clock::now();
        r.done = true;
        r.x.clear();
        r.x.shrink_to_fit();
        in_use -= static_cast<int64_t>(r.bytes);
        --inflight;
        ++completed;
        sl.req = -1;
    }

    void fail(int64_t id, Request& r, const cvm::Error& e) {
        r.status = e.code;
        r.err = e.what();
        r.done = true;
        r.t_done = Clock::now();
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        std::vector<std::pair<int64_t, int>> batch;
        while (true) {
            // retire finished evaluations
            for (size_t si = 0; si < slots.size(); ++si) {
                if (slots[si].req < 0) continue;
                const hipError_t q = hipEventQuery(slots[si].ev);
                if (q == hipSuccess) {
                    retire(static_cast<int>(si));
                    cv_done.notify_all();
                } else if (q != hipErrorNotReady) {
                    Request& r = reqs[slots[si].req];
                    r.status = CVM_E_HIP;
                    r.err = std::string("HIP error while evaluating a request: ") + hipGetErrorString(q);
                    r.done = true;
                    r.t_done = Clock::now();
                    in_use -= static_cast<int64_t>(r.bytes);
                    --inflight;
                    slots[si].req = -1;
                    cv_done.notify_all();
                }
            }
            // admit in arrival order
            while (!pending.empty()) {
                const int64_t id = pending.front();
                Request& r = reqs[id];
                // the smallest free slot that holds the request
                const SheetRef& sh = sheets[r.sheet];
                int use = -1;
                for (size_t si = 0; si < slots.size(); ++si) {
                    if (slots[si].req >= 0 || !fits(slots[si], sh, r.flags, r.K_dev == nullptr)) continue;
                    if (use < 0 || slots[si].bytes() < slots[use].bytes()) use = static_cast<int>(si);
                }
                if (use < 0) break;
                pending.pop_front();
                slots[use].req = id;             // reserved; launched with the pass's batch
                r.slot = use;
                in_use += static_cast<int64_t>(r.bytes);
                ++inflight;
                batch.emplace_back(id, use);
            }
            if (!batch.empty()) {
                launch_batch(batch);
                batch.clear();
            }
            if (stop && inflight == 0 && pending.empty()) break;
            if (inflight == 0) {
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
};

extern "C" {

int cvm_server_create(int device, int64_t hbm_budget, int max_inflight, cvm_server** out) {
    if (!out || max_inflight < 1 || max_inflight > 64 || hbm_budget <= 0) return CVM_E_INVALID;
    *out = nullptr;
    auto* s = new cvm_server();
    s->device = device;
    s->budget = hbm_budget;
    if (hipSetDevice(device) != hipSuccess) { delete s; return CVM_E_HIP; }
    s->slots.resize(max_inflight);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.ws.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            for (Slot& t : s->slots) {
                if (t.ws.stream) (void)hipStreamDestroy(t.ws.stream);
                if (t.ev) (void)hipEventDestroy(t.ev);
            }
            delete s;
            return CVM_E_HIP;
        }
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return CVM_OK;
}

void cvm_server_destroy(cvm_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (Slot& sl : s->slots) {
        if (sl.ws.stream) (void)hipStreamSynchronize(sl.ws.stream);
        if (sl.hx) (void)hipHostFree(sl.hx);
        if (sl.hF) (void)hipHostFree(sl.hF);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        sl.ws.scr.release(); sl.ws.fscr.release(); sl.ws.fold_gp.release(); sl.ws.fold_nd.release();
        sl.x.release(); sl.F.release(); sl.K.release();
        if (sl.ws.stream) (void)hipStreamDestroy(sl.ws.stream);
    }
    delete s;
}

const char* cvm_server_last_error(cvm_server* s) { return s ? s->err.c_str() : "null server"; }

int cvm_server_add_sheet(cvm_server* s, cvm_context* ctx, int* sheet_id) {
    if (!s || !ctx) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    cvm::Context& c = ctx->c;
    if (c.device != s->device) { s->err = "sheet context is on another device"; return CVM_E_INVALID; }
    if (!c.plan.ready) { s->err = "sheet has no assembly plan (cvm_init_assembly)"; return CVM_E_STATE; }
    if (s->inflight > 0 || !s->pending.empty()) {
        s->err = "sheets are registered while no request is pending or in flight";
        return CVM_E_STATE;
    }
    SheetRef r;
    r.ctx = ctx;
    r.nn = c.nn;
    r.nnz_union = static_cast<int64_t>(c.plan.uni_inner.size());
    r.nnz_face = static_cast<int64_t>(c.plan.face_inner.size());
    s->sheets.push_back(r);
    s->planned = false;
    if (sheet_id) *sheet_id = static_cast<int>(s->sheets.size()) - 1;
    return CVM_OK;
}

int cvm_server_request_bytes(cvm_server* s, int sheet, int flags, int with_K, int64_t* bytes) {
    if (!s || !bytes) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (sheet < 0 || sheet >= static_cast<int>(s->sheets.size())) { s->err = "bad sheet id"; return CVM_E_INVALID; }
    *bytes = static_cast<int64_t>(s->request_bytes(s->sheets[sheet], flags, with_K != 0));
    return CVM_OK;
}

int cvm_server_submit(cvm_server* s, int sheet, const double* x, int flags, double* F_host, double* K_dev,
                      int64_t* request_id) {
    if (!s || !x || !request_id) return CVM_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (sheet < 0 || sheet >= static_cast<int>(s->sheets.size())) { s->err = "bad sheet id"; return CVM_E_INVALID; }
    if (flags & CVM_DEVICE_PTRS) { s->err = "requests take host x and F (K may be a device buffer)"; return CVM_E_INVALID; }
    const SheetRef& sh = s->sheets[sheet];
    Request r;
    r.sheet = sheet;
    r.flags = flags | CVM_RESIDUAL;
    r.F_host = F_host;
    r.K_dev = K_dev;
    r.bytes = s->request_bytes(sh, r.flags, K_dev == nullptr);
    if (!s->planned) {
        if (s->inflight > 0 || !s->pending.empty()) { s->err = "server not idle while unplanned"; return CVM_E_STATE; }
        try {
            s->plan_slots();
        } catch (const cvm::Error& e) {
            s->err = e.what();
            return e.code;
        }
    }
    bool held = false;
    for (const Slot& sl : s->slots) held = held || cvm_server::fits(sl, sh, r.flags, K_dev == nullptr);
    if (static_cast<int64_t>(r.bytes) > s->budget || !held) {
        s->err = "request workspace (" + std::to_string(r.bytes) + " bytes) exceeds the HBM budget: no slot "
                 "provisioned within the budget holds it";
        return CVM_E_BUDGET;
    }
    r.x.assign(x, x + 3 * static_cast<size_t>(sh.nn));
    r.t_submit = Clock::now();
    const int64_t id = s->next_id++;
    s->reqs.emplace(id, std::move(r));
    s->pending.push_back(id);
    *request_id = id;
    lk.unlock();
    s->cv_work.notify_one();
    return CVM_OK;
}

// Blocks until request `id` is done; times (may be NULL) = queue ms (submit to
// launch), device ms (launch to completion seen), total ms. The request is
// forgotten afterwards.
int cvm_server_wait(cvm_server* s, int64_t id, double* times) {
    if (!s) return CVM_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(id) == s->reqs.end()) { s->err = "unknown request id"; return CVM_E_INVALID; }
    // another thread's wait on the same id may forget the request meanwhile
    s->cv_done.wait(lk, [&] {
        const auto f = s->reqs.find(id);
        return f == s->reqs.end() || f->second.done;
    });
    const auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { s->err = "request already waited for"; return CVM_E_INVALID; }
    Request& r = it->second;
    if (times) {
        const bool started = r.slot >= 0;
        times[0] = started ? ms_between(r.t_submit, r.t_start) : ms_between(r.t_submit, r.t_done);
        times[1] = started ? ms_between(r.t_start, r.t_done) : 0.0;
        times[2] = ms_between(r.t_submit, r.t_done);
    }
    const int st = r.status;
    if (st != CVM_OK) s->err = r.err;
    s->reqs.erase(id);
    return st;
}

int cvm_server_poll(cvm_server* s, int64_t id, int* done) {
    if (!s || !done) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { s->err = "unknown request id"; return CVM_E_INVALID; }
    *done = it->second.done ? 1 : 0;
    return CVM_OK;
}

int cvm_server_set_batching(cvm_server* s, int on) {
    if (!s) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->batching = on != 0;
    return CVM_OK;
}

int cvm_server_batching_stats(cvm_server* s, int64_t* out2) {
    if (!s || !out2) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    out2[0] = s->groups;
    out2[1] = s->grouped;
    return CVM_OK;
}

// completed requests, requests in flight, pending requests, workspace bytes allocated
int cvm_server_stats(cvm_server* s, int64_t* out4) {
    if (!s || !out4) return CVM_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    out4[0] = s->completed;
    out4[1] = s->inflight;
    out4[2] = static_cast<int64_t>(s->pending.size());
    out4[3] = static_cast<int64_t>(s->allocated());
    return CVM_OK;
}

}  // extern "C"
