// Fixed-radius neighbour search and crease indicator on gfx950.
//
// Neighbour lists must equal the reference's bit for bit (sets and order), so
// the distance test and the cell assignment use the reference's operations
// without fused multiply-add: floor((x - min) / rho) with IEEE division, and
// dist2 = dx*dx + dy*dy (core/include/curvami/FRNN.h:12-16, :63-72).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "cvm_internal.h"

namespace cvm {

// ---------------------------------------------------------------------------
// Crease indicator (core/include/curvami/Enrichment.h:23-92), evaluated once
// per point. The reference evaluates it per (pair, crease) inside the RKPM
// solve (GlobalRKPM.h:91-92); the value depends on the point only.
// ---------------------------------------------------------------------------
struct IndOut { double f, gx, gy, hxx, hyy, hxy; };

__device__ __forceinline__ IndOut indicator(double Px, double Py, const double* vx, const double* vy,
                                            int N, int ext_s, int ext_e, int closed) {
#pragma clang fp contract(off)
    IndOut o{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (N < 2) return o;
    const double inf = __builtin_inf();
    double d = inf;
    double Cbx = 0.0, Cby = 0.0, tb = 0.0, nbx = 0.0, nby = 0.0;
    int ib = 0;
    for (int i = 0; i < N - 1; ++i) {
        const double Ax = vx[i], Ay = vy[i];
        const double ux = vx[i + 1] - Ax, uy = vy[i + 1] - Ay;
        const double vsq = ux * ux + uy * uy;
        if (vsq == 0.0) continue;
        const double wx = Px - Ax, wy = Py - Ay;
        const double t = (wx * ux + wy * uy) / vsq;
        const double tmin = (i == 0 && ext_s) ? -inf : 0.0;
        const double tmax = (i == N - 2 && ext_e) ? inf : 1.0;
        const double tc = fmax(tmin, fmin(tmax, t));
        const double Cx = Ax + tc * ux, Cy = Ay + tc * uy;
        const double ex = Px - Cx, ey = Py - Cy;
        const double dsq = ex * ex + ey * ey;
        if (dsq < d * d) {
            d = sqrt(dsq);
            Cbx = Cx; Cby = Cy; ib = i; tb = tc;
            const double sv = sqrt(vsq);
            nbx = -uy / sv; nby = ux / sv;
        }
    }
    if (!(d > 1e-12)) return o;                       // on the crease: f = 0
    const double ux = (Px - Cbx) / d, uy = (Py - Cby) / d;
    const double dp = ux * nbx + uy * nby;
    const bool sc = (ib == 0 && tb == 0.0 && !ext_s && !closed);
    const bool ec = (ib == N - 2 && tb == 1.0 && !ext_e && !closed);
    if (sc || ec) {
        o.f = dp;
        o.gx = (nbx - o.f * ux) / d;
        o.gy = (nby - o.f * uy) / d;
        const double d2 = d * d;
        o.hxx = (3 * o.f * ux * ux - 2 * nbx * ux - o.f) / d2;
        o.hyy = (3 * o.f * uy * uy - 2 * nby * uy - o.f) / d2;
        o.hxy = (3 * o.f * ux * uy - (nbx * uy + ux * nby)) / d2;
    } else {
        o.f = (dp >= 0.0) ? 1.0 : -1.0;
    }
    return o;
}

// Per-GP indicator data for the creases of the GP's sheet, layout
// ind[(lc * 6 + k) * ng + g] with lc the crease's slot within its sheet
// (lane-coalesced). gsheet == null: one sheet holding every crease.
__global__ void k_indicator_gp(const double* __restrict__ gx, const double* __restrict__ gy, int ng,
                               const int* __restrict__ gsheet, const SheetDev* __restrict__ sheets,
                               const double* __restrict__ cv, const int* __restrict__ cinfo,
                               double* __restrict__ ind) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SheetDev& S = sheets[gsheet ? gsheet[g] : 0];
    const double px = gx[g], py = gy[g];
    for (int lc = 0; lc < S.nc; ++lc) {
        const int c = S.crease0 + lc;
        const int off = cinfo[5 * c], nv = cinfo[5 * c + 1];
        const IndOut o = indicator(px, py, cv + 2 * off, cv + 2 * off + nv, nv, cinfo[5 * c + 2],
                                   cinfo[5 * c + 3], cinfo[5 * c + 4]);
        double* dst = ind + static_cast<size_t>(lc) * 6 * ng + g;
        dst[0] = o.f; dst[ng] = o.gx; dst[2 * ng] = o.gy;
        dst[3 * ng] = o.hxx; dst[4 * ng] = o.hyy; dst[5 * ng] = o.hxy;
    }
}

// Node side of every crease of the node's sheet: +1 when f >= 0 (a node on the
// crease counts as the + side), -1 otherwise (GlobalRKPM.h:93); layout
// sign[lc * nn + i]. crange: per sheet (crease0, ncreases); nsheet == null: one sheet.
__global__ void k_node_sign(const double* __restrict__ X, const double* __restrict__ Y, int nn,
                            const int* __restrict__ nsheet, const int* __restrict__ crange,
                            const double* __restrict__ cv, const int* __restrict__ cinfo,
                            int8_t* __restrict__ sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int sh = nsheet ? nsheet[i] : 0;
    const int c0 = crange[2 * sh], nc = crange[2 * sh + 1];
    for (int lc = 0; lc < nc; ++lc) {
        const int c = c0 + lc;
        const int off = cinfo[5 * c], nv = cinfo[5 * c + 1];
        const IndOut o = indicator(X[i], Y[i], cv + 2 * off, cv + 2 * off + nv, nv, cinfo[5 * c + 2],
                                   cinfo[5 * c + 3], cinfo[5 * c + 4]);
        sign[static_cast<size_t>(lc) * nn + i] = (o.f >= 0.0) ? 1 : -1;
    }
}

static void upload_creases(Context& c) {
    if (!c.creases_dirty) return;
    std::vector<double> cv;
    std::vector<int> info;
    for (const auto& cr : c.creases) {
        const int off = static_cast<int>(cv.size() / 2);
        const int nv = static_cast<int>(cr.vx.size());
        // per crease: nv x values then nv y values at pair offset `off`
        std::vector<double> blk(2 * nv);
        for (int i = 0; i < nv; ++i) { blk[i] = cr.vx[i]; blk[nv + i] = cr.vy[i]; }
        cv.insert(cv.end(), blk.begin(), blk.end());
        info.insert(info.end(), {off, nv, cr.ext_s, cr.ext_e, cr.closed});
    }
    if (cv.empty()) cv.push_back(0.0);
    if (info.empty()) info.assign(5, 0);
    c.d_cv.upload(cv, c.stream);
    c.d_cinfo.upload(info, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    c.creases_dirty = false;
}

void launch_node_signs(Context& c) {
    sync_nodes(c);
    upload_creases(c);
    const int nc = static_cast<int>(c.creases.size());
    if (c.node_sign_creases == nc) return;        // unchanged nodes and creases
    c.node_sign_creases = nc;
    const int mc = c.max_sheet_creases();
    c.d_node_sign.alloc(std::max<size_t>(1, static_cast<size_t>(mc) * c.nn));
    if (mc == 0 || c.nn == 0) return;
    std::vector<int> crange;
    if (c.batched)
        for (const SheetSeg& s : c.sheets) crange.insert(crange.end(), {s.crease0, s.ncreases});
    else
        crange = {0, nc};
    if (crange != c.h_crange) {
        c.h_crange = crange;
        c.d_crange.upload(c.h_crange, c.stream);
    }
    const int slot = timing_begin(c, "node_sign");
    hipLaunchKernelGGL(k_node_sign, dim3((c.nn + 255) / 256), dim3(256), 0, c.stream, c.X.p, c.Y.p, c.nn,
                       c.batched ? c.d_nsheet.p : nullptr, c.d_crange.p, c.d_cv.p, c.d_cinfo.p, c.d_node_sign.p);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
}

// The crease range (first crease, count) of every sheet in the domain's
// SheetDev table must be the current one: creases can be added or replaced
// after the neighbour search, and the indicator and RKPM kernels index the
// per-crease arrays (sized for the current crease count) by it.
static void sync_sheet_creases(Context& c, Domain& d) {
    bool changed = false;
    for (size_t si = 0; si < d.h_sheets.size(); ++si) {
        const int nc = c.batched ? c.sheets.at(si).ncreases : static_cast<int>(c.creases.size());
        const int c0 = c.batched ? c.sheets.at(si).crease0 : 0;
        SheetDev& t = d.h_sheets[si];
        if (t.nc != nc || t.crease0 != c0) { t.nc = nc; t.crease0 = c0; changed = true; }
    }
    if (changed)
        d.d_sheets.upload(reinterpret_cast<const unsigned char*>(d.h_sheets.data()),
                          d.h_sheets.size() * sizeof(SheetDev), c.stream);
}

void launch_indicator_gp(Context& c, Domain& d) {
    upload_creases(c);
    sync_sheet_creases(c, d);
    const int nc = static_cast<int>(c.creases.size());
    if (d.ind_creases == nc) return;              // unchanged GPs and creases
    d.ind_creases = nc;
    const int mc = c.max_sheet_creases();
    d.ind.alloc(std::max<size_t>(1, static_cast<size_t>(mc) * 6 * d.ng));
    if (mc == 0 || d.ng == 0) return;
    const int slot = timing_begin(c, "indicator_gp");
    hipLaunchKernelGGL(k_indicator_gp, dim3((d.ng + 255) / 256), dim3(256), 0, c.stream, d.gx.p, d.gy.p, d.ng,
                       c.batched ? d.d_gsheet.p : nullptr, reinterpret_cast<const SheetDev*>(d.d_sheets.p),
                       c.d_cv.p, c.d_cinfo.p, d.ind.p);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
}

// ---------------------------------------------------------------------------
// Fixed-radius neighbours (FRNN.h:23-78). Nodes are binned by the key
// y*cols + x of their rho-sized cell, in ascending node index within a cell
// (the insertion order of the reference's per-cell vectors). A query scans
// the keys (qy+dy)*cols + (qx+dx), dx outer and dy inner, exactly as the
// reference does, including the wrap of an out-of-range x into the adjacent
// row's key; keys outside [0, nkeys) are empty, like a failed map lookup.
// ---------------------------------------------------------------------------
// FILL: write the lists at off (else count per GP). skip (may be null): the
// fill does nothing while *skip is set (a re-search whose counts changed).
template <bool FILL>
__global__ void k_frnn(const double* __restrict__ X, const double* __restrict__ Y,
                       const double* __restrict__ gx, const double* __restrict__ gy, int ng,
                       const int* __restrict__ gsheet, const SheetDev* __restrict__ sheets,
                       const int* __restrict__ cstart, const int* __restrict__ cnodes,
                       int* __restrict__ count, const int* __restrict__ off,
                       int* __restrict__ ids, double* __restrict__ s, const int* __restrict__ skip) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ng) return;
    if (FILL && skip && *skip) return;
    const SheetDev& gr = sheets[gsheet ? gsheet[q] : 0];
    const double px = gx[q], py = gy[q];
    const int qx = static_cast<int>(floor((px - gr.minx) / gr.cell));
    const int qy = static_cast<int>(floor((py - gr.miny) / gr.cell));
    const double rho = gr.rho, rho2 = gr.rho2;
    int n = 0;
    int o = FILL ? off[q] : 0;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            const long long key = static_cast<long long>(qy + dy) * gr.cols + (qx + dx);
            if (key < 0 || key >= gr.nkeys) continue;
            const int b = cstart[gr.koff + key], e = cstart[gr.koff + key + 1];
            for (int k = b; k < e; ++k) {
                const int ni = cnodes[k];
                const double ex = X[ni] - px, ey = Y[ni] - py;
                const double d2 = ex * ex + ey * ey;
                if (d2 <= rho2) {
                    if (FILL) {
                        ids[o] = ni;
                        s[o] = sqrt(d2) / rho;
                        ++o;
                    } else {
                        ++n;
                    }
                }
            }
        }
    }
    if (!FILL) count[q] = n;
}

__global__ void k_set_zero_int(int* p) { p[0] = 0; }

// flag[0] |= 1 where two offset arrays differ; clears flag[1] for the fill pass
__global__ void k_cmp_off(const int* __restrict__ a, const int* __restrict__ b, int n, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) flag[1] = 0;
    if (i < n && a[i] != b[i]) atomicOr(flag, 1);
}

__global__ void k_zero_int2(int* p) { p[0] = 0; p[1] = 0; }

// Interleaved copy of the pair lists for the moment pass (layout on Domain::il_row).
// CMP: the interleaved lists hold a previous search's result (same offsets);
// an entry that changes sets flag[1]; nothing is done when flag[0] is set
// (the offsets changed, so the in-place layout does not apply).
template <bool CMP>
__global__ __launch_bounds__(RKPM_BLOCK) void k_interleave(const int* __restrict__ off, const int* __restrict__ ids,
                                                           const double* __restrict__ s, int ng,
                                                           const int* __restrict__ row, int* __restrict__ il_ids,
                                                           double* __restrict__ il_s, uint8_t* __restrict__ lgp,
                                                           int* __restrict__ flag) {
    const int g = blockIdx.x * RKPM_BLOCK + threadIdx.x;
    if (g >= ng) return;
    if (CMP && flag[0]) return;
    const int b = off[g], n = off[g + 1] - b;
    const size_t base = static_cast<size_t>(row[blockIdx.x]) * RKPM_BLOCK + threadIdx.x;
    bool diff = false;
    for (int t = 0; t < n; ++t) {
        const size_t k = base + static_cast<size_t>(t) * RKPM_BLOCK;
        const int id = ids[b + t];
        const double sv = s[b + t];
        if (CMP) diff |= (il_ids[k] != id) || (__double_as_longlong(il_s[k]) != __double_as_longlong(sv));
        il_ids[k] = id;
        il_s[k] = sv;
        lgp[b + t] = static_cast<uint8_t>(threadIdx.x);
    }
    if (CMP && diff) atomicOr(flag + 1, 1);
}

// Per sheet: node cell grid with the bounds and column count of FRNN.h:30-42
// computed over that sheet's nodes; the sheets' key ranges are concatenated.
void launch_frnn(Context& c, Domain& d, double rho) {
    if (c.nn == 0) throw Error(CVM_E_STATE, "nodes not set");
    sync_nodes(c);
    sync_domain(c, d);
    std::vector<SheetSeg> segs;
    if (c.batched) {
        if (rho != 0.0) throw Error(CVM_E_INVALID, "batched context: pass rho = 0 (each sheet has its own radius)");
        segs = c.sheets;
    } else {
        if (!(rho > 0.0)) throw Error(CVM_E_INVALID, "rho must be > 0");
        SheetSeg one;
        one.node0 = 0; one.nn = c.nn; one.crease0 = 0; one.ncreases = static_cast<int>(c.creases.size());
        one.rho = rho;
        segs.push_back(one);
    }
    d.rho = segs[0].rho;
    std::vector<SheetDev> table(segs.size());
    std::vector<int> cstart(1, 0), cnodes(c.nn);
    long long koff = 0;
    int nfill = 0;
    for (size_t si = 0; si < segs.size(); ++si) {
        const SheetSeg& sg = segs[si];
        SheetDev& gr = table[si];
        gr.rho = sg.rho; gr.rho2 = sg.rho * sg.rho; gr.cell = sg.rho;
        gr.irho = 1.0 / sg.rho; gr.irho2 = 1.0 / (sg.rho * sg.rho);
        gr.nc = sg.ncreases; gr.crease0 = sg.crease0;
        if (sg.nn == 0) { gr.nkeys = 0; gr.koff = koff; continue; }
        double mnx = c.hX[sg.node0], mny = c.hY[sg.node0], mxx = mnx, mxy = mny;
        for (int i = sg.node0; i < sg.node0 + sg.nn; ++i) {
            mnx = std::min(mnx, c.hX[i]); mny = std::min(mny, c.hY[i]);
            mxx = std::max(mxx, c.hX[i]); mxy = std::max(mxy, c.hY[i]);
        }
        gr.minx = mnx; gr.miny = mny;
        gr.cols = static_cast<int>(std::ceil((mxx - mnx) / sg.rho)) + 1;
        std::vector<long long> key(sg.nn);
        long long maxkey = 0;
        for (int i = 0; i < sg.nn; ++i) {
            const int x = static_cast<int>(std::floor((c.hX[sg.node0 + i] - mnx) / sg.rho));
            const int y = static_cast<int>(std::floor((c.hY[sg.node0 + i] - mny) / sg.rho));
            key[i] = static_cast<long long>(y) * gr.cols + x;
            maxkey = std::max(maxkey, key[i]);
        }
        if (koff + maxkey + 1 > (1LL << 30)) throw Error(CVM_E_INVALID, "neighbour grid too large for this rho");
        gr.nkeys = static_cast<int>(maxkey + 1);
        gr.koff = koff;
        std::vector<int> cnt(gr.nkeys + 1, 0);
        for (int i = 0; i < sg.nn; ++i) cnt[key[i] + 1]++;
        for (int k = 0; k < gr.nkeys; ++k) cnt[k + 1] += cnt[k];
        for (int k = 1; k <= gr.nkeys; ++k) cstart.push_back(nfill + cnt[k]);
        std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < sg.nn; ++i) cnodes[nfill + fillp[key[i]]++] = sg.node0 + i;   // ascending per cell
        nfill += sg.nn;
        koff += gr.nkeys;
    }
    d.d_sheets.upload(reinterpret_cast<const unsigned char*>(table.data()), table.size() * sizeof(SheetDev), c.stream);
    d.h_sheets = table;
    d.off.alloc(d.ng + 1);
    d.h_off.assign(d.ng + 1, 0);
    d.ind_creases = -1;
    d.il_row.release(); d.il_ids.release(); d.il_s.release(); d.pair_lgp.release();
    if (d.ng == 0) {
        d.npairs = 0;
        CVM_HIP(hipMemsetAsync(d.off.p, 0, sizeof(int), c.stream));
        CVM_HIP(hipStreamSynchronize(c.stream));
        return;
    }
    const int* gsh = c.batched ? d.d_gsheet.p : nullptr;
    const SheetDev* tab = reinterpret_cast<const SheetDev*>(d.d_sheets.p);
    DBuf<int>& d_cstart = d.cell_start;
    DBuf<int>& d_cnodes = d.cell_nodes;
    DBuf<int>& d_count = d.cnt;
    d_cstart.upload(cstart, c.stream);
    d_cnodes.upload(cnodes, c.stream);
    d_count.alloc(d.ng);
    d.grid_nn = c.nn;
    d.grid_ng = d.ng;
    const dim3 grid((d.ng + 255) / 256), blk(256);
    int slot = timing_begin(c, "frnn_count");
    hipLaunchKernelGGL((k_frnn<false>), grid, blk, 0, c.stream, c.X.p, c.Y.p, d.gx.p, d.gy.p, d.ng, gsh, tab,
                       d_cstart.p, d_cnodes.p, d_count.p, nullptr, nullptr, nullptr, nullptr);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
    // off = [0, inclusive_scan(count)]
    hipLaunchKernelGGL(k_set_zero_int, dim3(1), dim3(1), 0, c.stream, d.off.p);
    size_t tmp_bytes = 0;
    CVM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, d_count.p, d.off.p + 1, d.ng, c.stream));
    DBuf<unsigned char>& tmp = d.scan_tmp;
    tmp.ensure(std::max<size_t>(tmp_bytes, 1));
    CVM_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tmp_bytes, d_count.p, d.off.p + 1, d.ng, c.stream));
    CVM_HIP(hipMemcpyAsync(d.h_off.data(), d.off.p, (d.ng + 1) * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    // off is int32: a domain with 2^31 or more pairs wraps the scan, which
    // shows as a decrease (every count is below 2^31)
    for (int g = 0; g < d.ng; ++g)
        if (d.h_off[g + 1] < d.h_off[g])
            throw Error(CVM_E_INVALID, "a domain has 2^31 or more neighbour pairs (int32 offsets)");
    d.npairs = d.h_off[d.ng];
    d.ids.alloc(std::max<int64_t>(d.npairs, 1));
    d.s.alloc(std::max<int64_t>(d.npairs, 1));
    slot = timing_begin(c, "frnn_fill");
    hipLaunchKernelGGL((k_frnn<true>), grid, blk, 0, c.stream, c.X.p, c.Y.p, d.gx.p, d.gy.p, d.ng, gsh, tab,
                       d_cstart.p, d_cnodes.p, nullptr, d.off.p, d.ids.p, d.s.p, nullptr);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
    // interleaved lists: a block's rows = its largest neighbour count
    const int nblk = (d.ng + RKPM_BLOCK - 1) / RKPM_BLOCK;
    std::vector<int> row(nblk + 1, 0);
    for (int b = 0; b < nblk; ++b) {
        int mx = 0;
        for (int g = b * RKPM_BLOCK; g < std::min(d.ng, (b + 1) * RKPM_BLOCK); ++g)
            mx = std::max(mx, d.h_off[g + 1] - d.h_off[g]);
        if (static_cast<int64_t>(row[b]) + mx > INT32_MAX / RKPM_BLOCK)
            throw Error(CVM_E_NO_MEMORY, "interleaved neighbour lists exceed 2^31 slots");
        row[b + 1] = row[b] + mx;
    }
    const size_t slots = std::max<size_t>(static_cast<size_t>(row[nblk]) * RKPM_BLOCK, 1);
    d.il_row.upload(row, c.stream);
    d.il_ids.alloc(slots);
    d.il_s.alloc(slots);
    d.pair_lgp.alloc(std::max<int64_t>(d.npairs, 1));
    static_assert(RKPM_BLOCK <= 256, "GP-in-block index is a byte");
    hipLaunchKernelGGL(k_interleave<false>, dim3(nblk), dim3(RKPM_BLOCK), 0, c.stream, d.off.p, d.ids.p, d.s.p, d.ng,
                       d.il_row.p, d.il_ids.p, d.il_s.p, d.pair_lgp.p, nullptr);
    CVM_HIP(hipGetLastError());
    d.h_ids.resize(d.npairs);
    if (d.npairs) CVM_HIP(hipMemcpyAsync(d.h_ids.data(), d.ids.p, d.npairs * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    d.have_sf = false;
}

// The search again on the device with the grid of the last one, for every
// searched domain of the context, all enqueued before one read-back: counts
// and their scan, compared with the current offsets (flag 0); with equal
// offsets the lists are rewritten in place and the interleaved copies redone,
// each entry compared with the previous interleaved copy, which holds the last
// search's lists in coalesced order (flag 1); with other offsets both passes
// do nothing (device-side test of flag 0) and the domain takes the full search
// (host grid, mirrors, interleave rows). No host loop over GPs or pairs.
// Returns, per domain, whether its lists changed.
std::vector<char> refresh_frnn_all(Context& c) {
    const size_t nd = c.doms.size();
    std::vector<char> changed(nd, 0), live(nd, 0);
    c.rf_flags.ensure(std::max<size_t>(2 * nd, 2));
    for (size_t k = 0; k < nd; ++k) {
        Domain& d = *c.doms[k];
        if (d.npairs < 0) continue;
        if (c.nodes_dirty || d.dirty || d.grid_nn != c.nn || d.grid_ng != d.ng || !d.cell_start.p) {
            launch_frnn(c, d, c.batched ? 0.0 : d.rho);
            changed[k] = 1;
            continue;
        }
        d.ind_creases = -1;             // the indicator is evaluated again, as in every RKPM solve
        if (d.ng == 0) continue;
        live[k] = 1;
        int* flag = c.rf_flags.p + 2 * k;
        const int* gsh = c.batched ? d.d_gsheet.p : nullptr;
        const SheetDev* tab = reinterpret_cast<const SheetDev*>(d.d_sheets.p);
        const dim3 grid((d.ng + 255) / 256), blk(256);
        d.off_new.ensure(d.ng + 1);
        int slot = timing_begin(c, "frnn_count");
        hipLaunchKernelGGL((k_frnn<false>), grid, blk, 0, c.stream, c.X.p, c.Y.p, d.gx.p, d.gy.p, d.ng, gsh, tab,
                           d.cell_start.p, d.cell_nodes.p, d.cnt.p, nullptr, nullptr, nullptr, nullptr);
        CVM_HIP(hipGetLastError());
        timing_end(c, slot);
        hipLaunchKernelGGL(k_zero_int2, dim3(1), dim3(1), 0, c.stream, flag);
        hipLaunchKernelGGL(k_set_zero_int, dim3(1), dim3(1), 0, c.stream, d.off_new.p);
        size_t tmp_bytes = 0;
        CVM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, d.cnt.p, d.off_new.p + 1, d.ng, c.stream));
        d.scan_tmp.ensure(std::max<size_t>(tmp_bytes, 1));
        CVM_HIP(hipcub::DeviceScan::InclusiveSum(d.scan_tmp.p, tmp_bytes, d.cnt.p, d.off_new.p + 1, d.ng,
                                                 c.stream));
        hipLaunchKernelGGL(k_cmp_off, dim3((d.ng + 256) / 256), dim3(256), 0, c.stream, d.off_new.p, d.off.p,
                           d.ng + 1, flag);
        CVM_HIP(hipGetLastError());
        slot = timing_begin(c, "frnn_fill");
        hipLaunchKernelGGL((k_frnn<true>), grid, blk, 0, c.stream, c.X.p, c.Y.p, d.gx.p, d.gy.p, d.ng, gsh, tab,
                           d.cell_start.p, d.cell_nodes.p, nullptr, d.off.p, d.ids.p, d.s.p, flag);
        CVM_HIP(hipGetLastError());
        timing_end(c, slot);
        const int nblk = (d.ng + RKPM_BLOCK - 1) / RKPM_BLOCK;
        slot = timing_begin(c, "interleave");
        hipLaunchKernelGGL(k_interleave<true>, dim3(nblk), dim3(RKPM_BLOCK), 0, c.stream, d.off.p, d.ids.p, d.s.p,
                           d.ng, d.il_row.p, d.il_ids.p, d.il_s.p, d.pair_lgp.p, flag);
        CVM_HIP(hipGetLastError());
        timing_end(c, slot);
    }
    std::vector<int> fl(2 * nd, 0);
    CVM_HIP(hipMemcpyAsync(fl.data(), c.rf_flags.p, 2 * nd * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    for (size_t k = 0; k < nd; ++k) {
        if (!live[k]) continue;
        Domain& d = *c.doms[k];
        if (fl[2 * k]) {                // other counts: the full search
            launch_frnn(c, d, c.batched ? 0.0 : d.rho);
            changed[k] = 1;
        } else if (fl[2 * k + 1]) {     // same counts, other nodes: the host mirror of the ids follows
            if (d.npairs)
                CVM_HIP(hipMemcpyAsync(d.h_ids.data(), d.ids.p, d.npairs * sizeof(int), hipMemcpyDeviceToHost,
                                       c.stream));
            CVM_HIP(hipStreamSynchronize(c.stream));
            d.have_sf = false;
            changed[k] = 1;
        }
    }
    return changed;
}

}  // namespace cvm
