// Fixed-radius neighbour search and crease indicator on gfx950.
//
// Neighbour lists must equal the reference's bit for bit (sets and order), so
// the distance test and the cell assignment use the reference's operations
// without fused multiply-add: floor((x - min) / rho) with IEEE division, and
// dist2 = dx*dx + dy*dy (core/include/curvami/FRNN.h:12-16, :63-72).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "cvm_internal.h"

namespace cvm {

// ---------------------------------------------------------------------------
// Crease indicator (core/include/curvami/Enrichment.h:23-92), evaluated once
// per point. The reference evaluates it per (pair, crease) inside the RKPM
// solve (GlobalRKPM.h:91-92); the value depends on the point only.
// ---------------------------------------------------------------------------
struct IndOut { double f, gx, gy, hxx, hyy, hxy; };

// Segments per culling group of a crease polyline. Per group (CGRP_WORDS
// doubles): bounding box x0, y0, x1, y1 of its vertices (infinite on the
// sides an extended end segment reaches) and its first vertex.
constexpr int CGRP_SEG = 16;
constexpr int CGRP_WORDS = 6;
constexpr int CINFO_WORDS = 7;      // per crease: offset, count, ext_s, ext_e, closed, first group, groups

// The reference's scan over all segments (Enrichment.h:35-60) visits every
// segment in order and keeps the first one of smallest distance. Here a group
// of segments is visited only when its bounding box lies within U + margin of
// P, U being the distance to the nearest group's first vertex (a point on the
// polyline, so the true minimum is <= U). A skipped segment is farther than
// U + margin, hence farther than every segment that can be accepted once the
// nearest segment has been, and the accepted chain among the visited segments
// is the same as in the full scan: the result is identical. The margin
// (1e-9 (1 + |P| + max |vertex|)) exceeds the rounding of the distance
// arithmetic by six orders of magnitude.
__device__ __forceinline__ IndOut indicator(double Px, double Py, const double* vx, const double* vy,
                                            int N, int ext_s, int ext_e, int closed,
                                            const double* __restrict__ grp, int ngrp, double vmax) {
#pragma clang fp contract(off)
    IndOut o{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (N < 2) return o;
    const double inf = __builtin_inf();
    double u2 = inf;
    for (int j = 0; j < ngrp; ++j) {
        const double ex = Px - grp[j * CGRP_WORDS + 4], ey = Py - grp[j * CGRP_WORDS + 5];
        u2 = fmin(u2, ex * ex + ey * ey);
    }
    const double reach = sqrt(u2) + 1e-9 * (1.0 + fabs(Px) + fabs(Py) + vmax);
    const double reach2 = reach * reach;
    double d = inf;
    double Cbx = 0.0, Cby = 0.0, tb = 0.0, nbx = 0.0, nby = 0.0;
    int ib = 0;
    for (int j = 0; j < ngrp; ++j) {
        const double* G = grp + j * CGRP_WORDS;
        const double bx = fmax(fmax(G[0] - Px, Px - G[2]), 0.0);
        const double by = fmax(fmax(G[1] - Py, Py - G[3]), 0.0);
        if (bx * bx + by * by > reach2) continue;
        const int i1 = min(N - 1, (j + 1) * CGRP_SEG);
        for (int i = j * CGRP_SEG; i < i1; ++i) {
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

__device__ __forceinline__ IndOut crease_indicator(double px, double py, int c, const double* __restrict__ cv,
                                                   const int* __restrict__ cinfo, const double* __restrict__ cgrp,
                                                   const double* __restrict__ cmax) {
    const int* ci = cinfo + CINFO_WORDS * c;
    const int off = ci[0], nv = ci[1];
    return indicator(px, py, cv + 2 * off, cv + 2 * off + nv, nv, ci[2], ci[3], ci[4],
                     cgrp + static_cast<size_t>(ci[5]) * CGRP_WORDS, ci[6], cmax[c]);
}

// Per-GP indicator data for the creases of the GP's sheet, layout
// ind[(lc * 6 + k) * ng + g] with lc the crease's slot within its sheet
// (lane-coalesced); every domain of the table in one launch.
struct IndDom {
    const double* gx; const double* gy; const int* gsheet; const SheetDev* sheets;
    double* ind; int ng; int blk0;
};

template <typename T>
__device__ __forceinline__ const T& table_entry(const T* __restrict__ doms, int nd, int b) {
    int lo = 0, hi = nd;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (doms[mid].blk0 <= b) lo = mid; else hi = mid;
    }
    return doms[lo];
}

__global__ __launch_bounds__(256) void k_indicator_gp(const IndDom* __restrict__ doms, int nd,
                                                      const double* __restrict__ cv, const int* __restrict__ cinfo,
                                                      const double* __restrict__ cgrp,
                                                      const double* __restrict__ cmax) {
    const IndDom& D = table_entry(doms, nd, blockIdx.x);
    const int g = (blockIdx.x - D.blk0) * 256 + threadIdx.x;
    const int ng = D.ng;
    if (g >= ng) return;
    const SheetDev& S = D.sheets[D.gsheet ? D.gsheet[g] : 0];
    const double px = D.gx[g], py = D.gy[g];
    for (int lc = 0; lc < S.nc; ++lc) {
        const IndOut o = crease_indicator(px, py, S.crease0 + lc, cv, cinfo, cgrp, cmax);
        double* dst = D.ind + static_cast<size_t>(lc) * 6 * ng + g;
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
                            const double* __restrict__ cgrp, const double* __restrict__ cmax,
                            int8_t* __restrict__ sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int sh = nsheet ? nsheet[i] : 0;
    const int c0 = crange[2 * sh], nc = crange[2 * sh + 1];
    for (int lc = 0; lc < nc; ++lc) {
        const IndOut o = crease_indicator(X[i], Y[i], c0 + lc, cv, cinfo, cgrp, cmax);
        sign[static_cast<size_t>(lc) * nn + i] = (o.f >= 0.0) ? 1 : -1;
    }
}

static void upload_creases(Context& c) {
    if (!c.creases_dirty) return;
    std::vector<double> cv, grp, cmax;
    std::vector<int> info;
    const double inf = std::numeric_limits<double>::infinity();
    for (const auto& cr : c.creases) {
        const int off = static_cast<int>(cv.size() / 2);
        const int nv = static_cast<int>(cr.vx.size());
        // per crease: nv x values then nv y values at pair offset `off`
        std::vector<double> blk(2 * nv);
        double vmax = 0.0;
        for (int i = 0; i < nv; ++i) {
            blk[i] = cr.vx[i]; blk[nv + i] = cr.vy[i];
            vmax = std::max(vmax, std::max(std::fabs(cr.vx[i]), std::fabs(cr.vy[i])));
        }
        cv.insert(cv.end(), blk.begin(), blk.end());
        const int g0 = static_cast<int>(grp.size() / CGRP_WORDS);
        const int nseg = std::max(0, nv - 1);
        const int ngrp = (nseg + CGRP_SEG - 1) / CGRP_SEG;
        for (int j = 0; j < ngrp; ++j) {
            const int i0 = j * CGRP_SEG, i1 = std::min(nseg, (j + 1) * CGRP_SEG);   // segments [i0, i1)
            double x0 = cr.vx[i0], x1 = x0, y0 = cr.vy[i0], y1 = y0;
            for (int i = i0; i <= i1; ++i) {
                x0 = std::min(x0, cr.vx[i]); x1 = std::max(x1, cr.vx[i]);
                y0 = std::min(y0, cr.vy[i]); y1 = std::max(y1, cr.vy[i]);
            }
            // an extended end segment is a ray: its group is never skipped
            if ((i0 == 0 && cr.ext_s) || (i1 == nseg && cr.ext_e)) { x0 = -inf; y0 = -inf; x1 = inf; y1 = inf; }
            grp.insert(grp.end(), {x0, y0, x1, y1, cr.vx[i0], cr.vy[i0]});
        }
        info.insert(info.end(), {off, nv, cr.ext_s, cr.ext_e, cr.closed, g0, ngrp});
        cmax.push_back(vmax);
    }
    if (cv.empty()) cv.push_back(0.0);
    if (grp.empty()) grp.assign(CGRP_WORDS, 0.0);
    if (cmax.empty()) cmax.push_back(0.0);
    if (info.empty()) info.assign(CINFO_WORDS, 0);
    c.d_cv.upload(cv, c.stream);
    c.d_cinfo.upload(info, c.stream);
    c.d_cgrp.upload(grp, c.stream);
    c.d_cmax.upload(cmax, c.stream);
    CVM_HIP(hipStreamSynchronize(c.stream));
    c.creases_dirty = false;
}

hipStream_t fork_to(Context& c, hipStream_t kst) {
    if (!kst || kst == c.stream) return c.stream;
    CVM_HIP(hipEventRecord(c.ev_fork, c.stream));
    CVM_HIP(hipStreamWaitEvent(kst, c.ev_fork, 0));
    return kst;
}

void launch_node_signs(Context& c, hipStream_t kst) {
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
    const hipStream_t ks = fork_to(c, kst);
    const int slot = timing_begin(c, "node_sign", ks);
    hipLaunchKernelGGL(k_node_sign, dim3((c.nn + 255) / 256), dim3(256), 0, ks, c.X.p, c.Y.p, c.nn,
                       c.batched ? c.d_nsheet.p : nullptr, c.d_crange.p, c.d_cv.p, c.d_cinfo.p, c.d_cgrp.p,
                       c.d_cmax.p, c.d_node_sign.p);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot, ks);
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

// Indicator values of the domains whose values are stale, one launch.
void launch_indicators(Context& c, const std::vector<Domain*>& ds, hipStream_t kst) {
    upload_creases(c);
    const int nc = static_cast<int>(c.creases.size());
    const int mc = c.max_sheet_creases();
    std::vector<IndDom> tab;
    int64_t nblocks = 0;
    for (Domain* dp : ds) {
        Domain& d = *dp;
        sync_sheet_creases(c, d);
        if (d.ind_creases == nc) continue;        // unchanged GPs and creases
        d.ind_creases = nc;
        d.ind.alloc(std::max<size_t>(1, static_cast<size_t>(mc) * 6 * d.ng));
        if (mc == 0 || d.ng == 0) continue;
        IndDom r{};
        r.gx = d.gx.p; r.gy = d.gy.p; r.gsheet = c.batched ? d.d_gsheet.p : nullptr;
        r.sheets = reinterpret_cast<const SheetDev*>(d.d_sheets.p);
        r.ind = d.ind.p; r.ng = d.ng; r.blk0 = static_cast<int>(nblocks);
        nblocks += (d.ng + 255) / 256;
        if (nblocks > INT32_MAX) throw Error(CVM_E_INVALID, "too many Gauss points for one indicator launch");
        tab.push_back(r);
    }
    if (tab.empty()) return;
    const size_t bytes = tab.size() * sizeof(IndDom);
    if (c.ind_tab_host.size() != bytes || std::memcmp(c.ind_tab_host.data(), tab.data(), bytes) != 0) {
        c.ind_tab_host.assign(reinterpret_cast<const unsigned char*>(tab.data()),
                              reinterpret_cast<const unsigned char*>(tab.data()) + bytes);
        c.ind_tab.upload(c.ind_tab_host, c.stream);
    }
    const hipStream_t ks = fork_to(c, kst);
    const int slot = timing_begin(c, "indicator_gp", ks);
    hipLaunchKernelGGL(k_indicator_gp, dim3(static_cast<unsigned>(nblocks)), dim3(256), 0, ks,
                       reinterpret_cast<const IndDom*>(c.ind_tab.p), static_cast<int>(tab.size()), c.d_cv.p,
                       c.d_cinfo.p, c.d_cgrp.p, c.d_cmax.p);
    CVM_HIP(hipGetLastError());
    timing_end(c, slot, ks);
}

// ---------------------------------------------------------------------------
// Fixed-radius neighbours (FRNN.h:23-78). Nodes are binned by the key
// y*cols + x of their rho-sized cell, in ascending node index within a cell
// (the insertion order of the reference's per-cell vectors). A query scans
// the keys (qy+dy)*cols + (qx+dx), dx outer and dy inner, exactly as the
// reference does, including the wrap of an out-of-range x into the adjacent
// row's key; keys outside [0, nkeys) are empty, like a failed map lookup.
// ---------------------------------------------------------------------------
// FILL: write the lists at off (else count per GP).
template <bool FILL>
__global__ void k_frnn(const double* __restrict__ X, const double* __restrict__ Y,
                       const double* __restrict__ gx, const double* __restrict__ gy, int ng,
                       const int* __restrict__ gsheet, const SheetDev* __restrict__ sheets,
                       const int* __restrict__ cstart, const int* __restrict__ cnodes,
                       int* __restrict__ count, const int* __restrict__ off,
                       int* __restrict__ ids, double* __restrict__ s) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ng) return;
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

// Interleaved copy of the pair lists for the moment pass (layout on Domain::il_row).
__global__ __launch_bounds__(RKPM_BLOCK) void k_interleave(const int* __restrict__ off, const int* __restrict__ ids,
                                                           const double* __restrict__ s, int ng,
                                                           const int* __restrict__ row, int* __restrict__ il_ids,
                                                           double* __restrict__ il_s, uint8_t* __restrict__ lgp) {
    const int g = blockIdx.x * RKPM_BLOCK + threadIdx.x;
    if (g >= ng) return;
    const int b = off[g], n = off[g + 1] - b;
    const size_t base = static_cast<size_t>(row[blockIdx.x]) * RKPM_BLOCK + threadIdx.x;
    for (int t = 0; t < n; ++t) {
        const size_t k = base + static_cast<size_t>(t) * RKPM_BLOCK;
        il_ids[k] = ids[b + t];
        il_s[k] = s[b + t];
        lgp[b + t] = static_cast<uint8_t>(threadIdx.x);
    }
}

// The search again for every GP of the table's domains, each list compared
// entry by entry (node and the bits of s) with the resident interleaved copy,
// which holds the last search's lists; *flag = 1 for a domain where any list
// differs in length or content. 256 threads = 4 waves = 4 interleave blocks.
struct VerDom {
    const double* gx; const double* gy; const int* gsheet; const SheetDev* sheets;
    const int* cstart; const int* cnodes; const int* off; const int* il_row;
    const int* il_ids; const double* il_s; int* flag; int ng; int blk0;
};

__global__ __launch_bounds__(256) void k_frnn_verify(const double* __restrict__ X, const double* __restrict__ Y,
                                                     const VerDom* __restrict__ doms, int nd) {
#pragma clang fp contract(off)
    static_assert(256 % RKPM_BLOCK == 0, "a block covers whole interleave blocks");
    const VerDom& D = table_entry(doms, nd, blockIdx.x);
    const int q = (blockIdx.x - D.blk0) * 256 + threadIdx.x;
    if (q >= D.ng) return;
    const SheetDev& gr = D.sheets[D.gsheet ? D.gsheet[q] : 0];
    const double px = D.gx[q], py = D.gy[q];
    const int qx = static_cast<int>(floor((px - gr.minx) / gr.cell));
    const int qy = static_cast<int>(floor((py - gr.miny) / gr.cell));
    const double rho = gr.rho, rho2 = gr.rho2;
    const int old = D.off[q + 1] - D.off[q];
    const size_t base = static_cast<size_t>(D.il_row[q / RKPM_BLOCK]) * RKPM_BLOCK + (q % RKPM_BLOCK);
    const int* __restrict__ lid = D.il_ids + base;
    const double* __restrict__ ls = D.il_s + base;
    const int* __restrict__ cstart = D.cstart;
    const int* __restrict__ cnodes = D.cnodes;
    // the expected entries of the next four hits in a shift register (the entry
    // four hits ahead is loaded as one is consumed, so a hit does not wait for a
    // load issued at the previous one); the candidate's node index one candidate
    // ahead
    constexpr int LA = 4;
    int eid[LA];
    long long es[LA];
#pragma unroll
    for (int j = 0; j < LA; ++j) {
        eid[j] = j < old ? lid[static_cast<size_t>(j) * RKPM_BLOCK] : -1;
        es[j] = j < old ? __double_as_longlong(ls[static_cast<size_t>(j) * RKPM_BLOCK]) : 0;
    }
    int n = 0;
    bool diff = false;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            const long long key = static_cast<long long>(qy + dy) * gr.cols + (qx + dx);
            if (key < 0 || key >= gr.nkeys) continue;
            const int b = cstart[gr.koff + key], e = cstart[gr.koff + key + 1];
            int nn = b < e ? cnodes[b] : 0;
            for (int k = b; k < e; ++k) {
                const int ni = nn;
                if (k + 1 < e) nn = cnodes[k + 1];
                const double ex = X[ni] - px, ey = Y[ni] - py;
                const double d2 = ex * ex + ey * ey;
                if (d2 <= rho2) {
                    if (n < old) {      // slot n of this lane exists: n < old <= the block's row count
                        const double sv = sqrt(d2) / rho;
                        diff |= (eid[0] != ni) || (es[0] != __double_as_longlong(sv));
#pragma unroll
                        for (int j = 0; j + 1 < LA; ++j) { eid[j] = eid[j + 1]; es[j] = es[j + 1]; }
                        if (n + LA < old) {
                            const size_t k2 = static_cast<size_t>(n + LA) * RKPM_BLOCK;
                            eid[LA - 1] = lid[k2];
                            es[LA - 1] = __double_as_longlong(ls[k2]);
                        }
                    }
                    ++n;
                }
            }
        }
    }
    if (diff || n != old) atomicOr(D.flag, 1);
}

__global__ void k_zero_flags(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
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
                       d_cstart.p, d_cnodes.p, d_count.p, nullptr, nullptr, nullptr);
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
                       d_cstart.p, d_cnodes.p, nullptr, d.off.p, d.ids.p, d.s.p);
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
    hipLaunchKernelGGL(k_interleave, dim3(nblk), dim3(RKPM_BLOCK), 0, c.stream, d.off.p, d.ids.p, d.s.p, d.ng,
                       d.il_row.p, d.il_ids.p, d.il_s.p, d.pair_lgp.p);
    CVM_HIP(hipGetLastError());
    d.h_ids.resize(d.npairs);
    if (d.npairs) CVM_HIP(hipMemcpyAsync(d.h_ids.data(), d.ids.p, d.npairs * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamSynchronize(c.stream));
    d.have_sf = false;
}

// The search again on the device with the grid of the last one, for every
// searched domain of the context, in one launch before one read-back: every
// list is computed again and compared with the resident one (k_frnn_verify).
// A domain whose lists all match keeps them, which are then the new search's
// result; a domain with any difference takes the full search (host grid,
// counts, offsets, lists, interleaved copies and mirrors). No host loop over
// GPs or pairs. Returns, per domain, whether its lists changed.
std::vector<char> refresh_frnn_all(Context& c, const std::function<void(hipStream_t)>& beside,
                                   const std::function<void()>& joined) {
    const size_t nd = c.doms.size();
    std::vector<char> changed(nd, 0), live(nd, 0);
    c.node_sign_creases = -1;          // node sides are evaluated again, here or with the next shape functions
    c.rf_flags.ensure(std::max<size_t>(nd, 1));
    std::vector<VerDom> tab;
    int64_t nblocks = 0;
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
        if (d.il_row.n != static_cast<size_t>((d.ng + RKPM_BLOCK - 1) / RKPM_BLOCK + 1))
            throw Error(CVM_E_STATE, "interleaved neighbour lists missing for this domain");
        live[k] = 1;
        VerDom v{};
        v.gx = d.gx.p; v.gy = d.gy.p; v.gsheet = c.batched ? d.d_gsheet.p : nullptr;
        v.sheets = reinterpret_cast<const SheetDev*>(d.d_sheets.p);
        v.cstart = d.cell_start.p; v.cnodes = d.cell_nodes.p; v.off = d.off.p; v.il_row = d.il_row.p;
        v.il_ids = d.il_ids.p; v.il_s = d.il_s.p; v.flag = c.rf_flags.p + k; v.ng = d.ng;
        v.blk0 = static_cast<int>(nblocks);
        nblocks += (d.ng + 255) / 256;
        if (nblocks > INT32_MAX) throw Error(CVM_E_INVALID, "too many Gauss points for one neighbour launch");
        tab.push_back(v);
    }
    if (tab.empty()) return changed;
    const size_t bytes = tab.size() * sizeof(VerDom);
    if (c.rf_tab_host.size() != bytes || std::memcmp(c.rf_tab_host.data(), tab.data(), bytes) != 0) {
        c.rf_tab_host.assign(reinterpret_cast<const unsigned char*>(tab.data()),
                             reinterpret_cast<const unsigned char*>(tab.data()) + bytes);
        c.rf_tab.upload(c.rf_tab_host, c.stream);
    }
    // The verified domains' crease indicators, which depend on neither the
    // lists nor their verification, run on the side stream beside the
    // verification pass, and the node sides on the context's stream ahead of
    // it; side work enqueued by `beside` follows the indicators and (through
    // its fork) the node sides. The context's stream joins the side stream
    // before the flags are read. (A domain whose lists then differ takes the
    // full search, which marks its indicators stale again.)
    const hipStream_t sd = side_stream(c);
    std::vector<Domain*> ind_ds;
    for (size_t k = 0; k < nd; ++k)
        if (live[k]) ind_ds.push_back(c.doms[k]);
    launch_indicators(c, ind_ds, sd);
    launch_node_signs(c);
    if (beside) {
        try {
            beside(sd);
        } catch (...) {
            (void)hipStreamSynchronize(sd);   // nothing of the side stream outlives the call
            throw;
        }
    }
    CVM_HIP(hipEventRecord(c.ev_join, sd));
    hipLaunchKernelGGL(k_zero_flags, dim3((static_cast<unsigned>(nd) + 255) / 256), dim3(256), 0, c.stream,
                       c.rf_flags.p, static_cast<int>(nd));
    CVM_HIP(hipGetLastError());
    const int slot = timing_begin(c, "frnn_verify");
    hipLaunchKernelGGL(k_frnn_verify, dim3(static_cast<unsigned>(nblocks)), dim3(256), 0, c.stream, c.X.p, c.Y.p,
                       reinterpret_cast<const VerDom*>(c.rf_tab.p), static_cast<int>(tab.size()));
    CVM_HIP(hipGetLastError());
    timing_end(c, slot);
    std::vector<int> fl(nd, 0);
    CVM_HIP(hipMemcpyAsync(fl.data(), c.rf_flags.p, nd * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CVM_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
    if (joined) joined();
    CVM_HIP(hipStreamSynchronize(c.stream));
    for (size_t k = 0; k < nd; ++k) {
        if (!live[k] || !fl[k]) continue;
        launch_frnn(c, *c.doms[k], c.batched ? 0.0 : c.doms[k]->rho);
        changed[k] = 1;
    }
    return changed;
}

}  // namespace cvm
