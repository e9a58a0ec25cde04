// Assembly plan: the block-upper sparsity pattern of K (face pairs united with
// the elastic-fold pairs), the tiling of the face Gauss points, and the
// precomputed slot maps that make the device scatter deterministic.
//
// Pattern semantics follow NonLinearAssembler::initSparsityPattern
// (plate/Assembly.h:59-153): node block (lo, hi) with lo <= hi is present when
// some face GP has both nodes as neighbours; column 3*hi + c of the CSC holds,
// for each such lo in ascending order, the three rows 3*lo + r. The fold blocks
// are the upper-filtered triplets of Constraints::evaluateElasticFold
// (plate/Constraints.h:420-433) over every active crease GP, and
// Meshless::assembleSystem adds them as K + K_fold, i.e. over the union
// (plate/Meshless.h:118-126).
//
// Device scatter (no colliding atomics, run-to-run bit-identical):
//  * tile t = a run of consecutive face GPs whose neighbour union has at most
//    `union_max` nodes; every local pair of every GP of the tile is mapped to a
//    "touched" slot of the tile (pairs of union nodes that some GP of the tile
//    couples). The tile kernel accumulates its GPs in order into those slots in
//    LDS and writes them to a scratch row;
//  * for every K block, the list of scratch rows that hold a contribution,
//    ascending in tile (hence GP) order, is the segment the reduction kernel
//    sums, followed by the fold contributions in crease-GP order.
#include <algorithm>
#include <numeric>

#include "cvm_internal.h"

namespace cvm {

namespace {

Domain* face_domain(Context& c) {
    Domain* f = nullptr;
    for (Domain* d : c.doms)
        if (d->kind == CVM_DOMAIN_FACE) {
            if (f) throw Error(CVM_E_INVALID, "more than one face domain");
            f = d;
        }
    if (!f) throw Error(CVM_E_STATE, "no face domain");
    return f;
}


}  // namespace

void build_plan(Context& c) {
    Plan& P = c.plan;
    P = Plan();   // release old buffers
    Domain& fd = *face_domain(c);
    if (fd.npairs < 0 || !fd.have_sf) throw Error(CVM_E_STATE, "face shape functions not computed");
    const int nn = c.nn;
    P.nn = nn;
    P.tile_max_gps = c.tile_gps;
    P.union_max = c.union_max;
    const int ng = fd.ng;
    const std::vector<int>& off = fd.h_off;
    const std::vector<int>& ids = fd.h_ids;
    const std::vector<int>& gsh = fd.h_gsheet;
    int nmax = 0;
    for (int g = 0; g < ng; ++g) nmax = std::max(nmax, off[g + 1] - off[g]);
    if (nmax > kTileUnion) throw Error(CVM_E_INVALID, "a Gauss point has more than 255 neighbours");
    P.max_nbr = nmax;
    // default union cap: 64 up to 47 neighbours (4 x 4 Gauss cells at rho ~ 3.75 h),
    // proportionally more for larger supports
    const int auto_union = std::max(64, (nmax * 4 + 2) / 3);
    P.union_max = std::max(std::min(c.union_max > 0 ? c.union_max : auto_union, kTileUnion), nmax);

    // ---- tiles and touched pairs (tile size shrinks until the LDS fits) -----
    std::vector<int> tile_gp, tile_uoff, unodes;
    std::vector<int64_t> tile_toff;
    std::vector<int> tile_nd;       // per tile: paired slots (pairs of slots)
    std::vector<uint8_t> uloc(ids.size());
    std::vector<uint16_t> tpairs;
    std::vector<int> stamp(nn, -1), where(nn, -1);
    for (int tg = std::min(P.tile_max_gps, kTileGps);; tg = std::max(1, tg / 2)) {
        tile_gp.assign(1, 0); tile_uoff.assign(1, 0); unodes.clear();
        std::fill(stamp.begin(), stamp.end(), -1);
        // (1) tiles: greedy runs of consecutive GPs (sequential, O(pairs))
        int t = 0, g = 0;
        std::vector<int> cur;
        while (g < ng) {
            cur.clear();
            const int g0 = g;
            while (g < ng && g - g0 < tg) {
                // tiles never straddle two sheets of a batched context, so each
                // sheet's tiles (and sums) are those it has on its own
                if (!gsh.empty() && gsh[g] != gsh[g0]) break;
                int add = 0;
                for (int k = off[g]; k < off[g + 1]; ++k)
                    if (stamp[ids[k]] != t) { stamp[ids[k]] = t; ++add; cur.push_back(ids[k]); }
                if (static_cast<int>(cur.size()) > P.union_max && g > g0) {
                    for (int k = 0; k < add; ++k) { stamp[cur.back()] = -1; cur.pop_back(); }
                    break;
                }
                ++g;
            }
            std::sort(cur.begin(), cur.end());
            for (size_t k = 0; k < cur.size(); ++k) where[cur[k]] = static_cast<int>(k);
            for (int q = g0; q < g; ++q)
                for (int k = off[q]; k < off[q + 1]; ++k) uloc[k] = static_cast<uint8_t>(where[ids[k]]);
            unodes.insert(unodes.end(), cur.begin(), cur.end());
            tile_gp.push_back(g);
            tile_uoff.push_back(static_cast<int>(unodes.size()));
            ++t;
        }
        P.ntiles = t;
        // (2) touched pairs of each tile (parallel over tiles): (ua, ub) coupled by at
        // least one GP; sorted by the number of coupling GPs (descending), then by the
        // presence mask, so that the lanes of a wave walk equally long, mostly
        // identical GP sequences. Consecutive pairs with the same mask and the same
        // lower node are paired (one thread sums both blocks and reads their shared
        // records once per GP): the tile's list holds its paired slots first, two
        // by two, then the single ones.
        std::vector<std::vector<uint16_t>> tl(P.ntiles);
        tile_nd.assign(P.ntiles, 0);
        int max_touched = 0, max_union = 0, rec_cap = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : max_touched, max_union, rec_cap)
        for (int ti = 0; ti < P.ntiles; ++ti) {
            const int q0 = tile_gp[ti], q1 = tile_gp[ti + 1];
            const int U = tile_uoff[ti + 1] - tile_uoff[ti];
            std::vector<uint32_t> pres(U, 0u);
            for (int q = q0; q < q1; ++q)
                for (int k = off[q]; k < off[q + 1]; ++k) pres[uloc[k]] |= 1u << (q - q0);
            std::vector<std::pair<uint32_t, uint16_t>> tlist;
            for (int a = 0; a < U; ++a)
                for (int bb = a; bb < U; ++bb) {
                    const uint32_t m = pres[a] & pres[bb];
                    if (m) tlist.push_back({m, static_cast<uint16_t>((a << 8) | bb)});
                }
            std::stable_sort(tlist.begin(), tlist.end(), [](const auto& x, const auto& y) {
                const int cx = __builtin_popcount(x.first), cy = __builtin_popcount(y.first);
                if (cx != cy) return cx > cy;
                return x.first < y.first;
            });
            std::vector<uint16_t> sgl;
            tl[ti].clear();
            tl[ti].reserve(tlist.size());
            // A run of slots with the same mask and lower node pairs its first half
            // with its second half (slot i with slot i + L/2): the lanes that take
            // consecutive pairs then read the upper nodes' records one record apart
            // (18 dwords: distinct LDS banks for 16 lanes) instead of two apart (36
            // dwords: two lanes per bank). An odd run leaves its last slot single.
            for (size_t k = 0; k < tlist.size();) {
                size_t e = k + 1;
                while (e < tlist.size() && tlist[e].first == tlist[k].first &&
                       (tlist[e].second >> 8) == (tlist[k].second >> 8))
                    ++e;
                const size_t half = (e - k) / 2;
                for (size_t i = 0; i < half; ++i) {
                    tl[ti].push_back(tlist[k + i].second);
                    tl[ti].push_back(tlist[k + half + i].second);
                }
                if ((e - k) & 1) sgl.push_back(tlist[e - 1].second);
                k = e;
            }
            tile_nd[ti] = static_cast<int>(tl[ti].size() / 2);
            tl[ti].insert(tl[ti].end(), sgl.begin(), sgl.end());
            max_touched = std::max(max_touched, static_cast<int>(tlist.size()));
            max_union = std::max(max_union, U);
            rec_cap = std::max(rec_cap, off[q1] - off[q0]);
        }
        tile_toff.assign(P.ntiles + 1, 0);
        for (int ti = 0; ti < P.ntiles; ++ti) tile_toff[ti + 1] = tile_toff[ti] + static_cast<int64_t>(tl[ti].size());
        tpairs.resize(tile_toff[P.ntiles]);
#pragma omp parallel for schedule(static)
        for (int ti = 0; ti < P.ntiles; ++ti)
            std::copy(tl[ti].begin(), tl[ti].end(), tpairs.begin() + tile_toff[ti]);
        P.max_touched = max_touched;
        P.max_union = max_union;
        P.rec_cap = std::max(rec_cap, 1);
        P.tile_max_gps = tg;
        if (tile_lds_bytes(P.rec_cap, P.max_union, false, P.max_nbr > 64) <= c.lds_limit) break;
        if (tg == 1) throw Error(CVM_E_INVALID, "assembly tile of one Gauss point exceeds the LDS limit");
    }
    // per face pair: the rank of its node among the GP's neighbours in union
    // order (union indices are distinct within a GP and below 256, and so are
    // the neighbour positions: one 16-bit key each)
    std::vector<uint8_t> urank(ids.size());
#pragma omp parallel for schedule(static, 1024)
    for (int g = 0; g < ng; ++g) {
        const int b = off[g], n = off[g + 1] - b;
        uint16_t key[kTileUnion];
        for (int k = 0; k < n; ++k) key[k] = static_cast<uint16_t>((uloc[b + k] << 8) | k);
        std::sort(key, key + n);
        for (int r = 0; r < n; ++r) urank[b + (key[r] & 0xff)] = static_cast<uint8_t>(r);
    }
    // face blocks: for each column node hi, the nodes lo <= hi that share a face GP
    // with it (parallel over columns through the node -> GP incidence)
    std::vector<std::vector<int>> cols(nn);     // face + fold rows per column
    {
        std::vector<int> ncnt(nn + 1, 0);
        for (size_t k = 0; k < ids.size(); ++k) ncnt[ids[k] + 1]++;
        for (int i = 0; i < nn; ++i) ncnt[i + 1] += ncnt[i];
        std::vector<int> ngp_of(ids.size());
        {
            std::vector<int> fillp(ncnt.begin(), ncnt.end() - 1);
            for (int q = 0; q < ng; ++q)
                for (int k = off[q]; k < off[q + 1]; ++k) ngp_of[fillp[ids[k]]++] = q;
        }
#pragma omp parallel
        {
            std::vector<int> mark(nn, -1);
#pragma omp for schedule(dynamic, 256)
            for (int hi = 0; hi < nn; ++hi) {
                auto& v = cols[hi];
                for (int a = ncnt[hi]; a < ncnt[hi + 1]; ++a) {
                    const int q = ngp_of[a];
                    for (int k = off[q]; k < off[q + 1]; ++k) {
                        const int lo = ids[k];
                        if (lo <= hi && mark[lo] != hi) { mark[lo] = hi; v.push_back(lo); }
                    }
                }
            }
        }
    }
    P.nscratch = static_cast<int64_t>(tpairs.size());
    P.nfscratch = static_cast<int64_t>(unodes.size());
    // scratch rows are addressed by uint32 slot_dst, union entries by int fc_idx
    if (P.nscratch > 0xffffffffLL) throw Error(CVM_E_INVALID, "assembly scratch exceeds 2^32 rows");
    if (P.nfscratch > 0x7fffffffLL) throw Error(CVM_E_INVALID, "assembly tiles hold more than 2^31 union entries");

    // face pattern (per column: sorted unique rows)
    std::vector<std::vector<int>> face_rows(nn);
#pragma omp parallel for schedule(dynamic, 256)
    for (int j = 0; j < nn; ++j) {
        auto& v = cols[j];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        face_rows[j] = v;
    }
    P.face_rows = face_rows;

    // ---- fold crease GPs -----------------------------------------------------
    struct FGP { int fold, local, left_b, left_n, right_b, right_n; std::vector<int> nodes; };
    std::vector<FGP> fgps;
    for (size_t f = 0; f < c.folds.size(); ++f) {
        const Fold& fo = c.folds[f];
        if (fo.rigid) continue;             // rigid hinges are constraint rows, not K terms
        Domain& L = *c.doms[fo.left];
        Domain& R = *c.doms[fo.right];
        if (L.npairs < 0 || R.npairs < 0 || !L.have_sf || !R.have_sf)
            throw Error(CVM_E_STATE, "fold domains need shape functions before the assembly plan");
        for (int i = 0; i < L.ng; ++i) {
            if (!fo.active.empty() && !fo.active[i]) continue;
            FGP q;
            q.fold = static_cast<int>(f); q.local = i;
            q.left_b = L.h_off[i]; q.left_n = L.h_off[i + 1] - L.h_off[i];
            q.right_b = R.h_off[i]; q.right_n = R.h_off[i + 1] - R.h_off[i];
            for (int k = 0; k < q.left_n; ++k) q.nodes.push_back(L.h_ids[q.left_b + k]);
            for (int k = 0; k < q.right_n; ++k) q.nodes.push_back(R.h_ids[q.right_b + k]);
            std::sort(q.nodes.begin(), q.nodes.end());
            q.nodes.erase(std::unique(q.nodes.begin(), q.nodes.end()), q.nodes.end());
            for (size_t a = 0; a < q.nodes.size(); ++a)
                for (size_t bb = a; bb < q.nodes.size(); ++bb) cols[q.nodes[bb]].push_back(q.nodes[a]);
            fgps.push_back(std::move(q));
        }
    }
    P.nfgp = static_cast<int>(fgps.size());

    // union pattern
    P.col_ptr.assign(nn + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int j = 0; j < nn; ++j) {
        auto& v = cols[j];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    for (int j = 0; j < nn; ++j) P.col_ptr[j + 1] = P.col_ptr[j] + static_cast<int>(cols[j].size());
    P.nblocks = P.col_ptr[nn];
    P.col_rows.resize(P.nblocks);
    P.blk_face.assign(P.nblocks, 0);
    std::vector<int64_t> base_u(P.nblocks), base_f(P.nblocks, -1);
    std::vector<int> stride_u(P.nblocks), stride_f(P.nblocks, 0);
    P.face_outer.assign(3 * nn + 1, 0);
    P.uni_outer.assign(3 * nn + 1, 0);
    {
        int64_t fo = 0, uo = 0;
        for (int j = 0; j < nn; ++j) {
            const auto& ur = cols[j];
            const auto& fr = face_rows[j];
            size_t fpos = 0;
            for (size_t k = 0; k < ur.size(); ++k) {
                const int64_t blk = P.col_ptr[j] + static_cast<int64_t>(k);
                P.col_rows[blk] = ur[k];
                base_u[blk] = uo + 3 * static_cast<int64_t>(k);
                stride_u[blk] = 3 * static_cast<int>(ur.size());
                while (fpos < fr.size() && fr[fpos] < ur[k]) ++fpos;
                if (fpos < fr.size() && fr[fpos] == ur[k]) {
                    P.blk_face[blk] = 1;
                    base_f[blk] = fo + 3 * static_cast<int64_t>(fpos);
                    stride_f[blk] = 3 * static_cast<int>(fr.size());
                }
            }
            for (int cc = 0; cc < 3; ++cc) {
                for (int r : fr)
                    for (int rr = 0; rr < 3; ++rr) P.face_inner.push_back(3 * r + rr);
                fo += 3 * static_cast<int64_t>(fr.size());
                P.face_outer[3 * j + cc + 1] = static_cast<int>(fo);
                for (int r : ur)
                    for (int rr = 0; rr < 3; ++rr) P.uni_inner.push_back(3 * r + rr);
                uo += 3 * static_cast<int64_t>(ur.size());
                P.uni_outer[3 * j + cc + 1] = static_cast<int>(uo);
            }
        }
        if (uo > 0x7fffffffLL) throw Error(CVM_E_INVALID, "K has more than 2^31 stored values (int32 CSC)");
    }
    auto block_of = [&](int lo, int hi) -> int64_t {
        const int* b = P.col_rows.data() + P.col_ptr[hi];
        const int* e = P.col_rows.data() + P.col_ptr[hi + 1];
        const int* it = std::lower_bound(b, e, lo);
        if (it == e || *it != lo) return -1;
        return static_cast<int64_t>(it - P.col_rows.data());
    };

    // ---- K contributors: face tiles then folds ---------------------------------
    std::vector<int64_t> kc_cnt(P.nblocks + 1, 0);
    std::vector<int64_t> tblk(tpairs.size());
    bool miss = false;
#pragma omp parallel for schedule(dynamic, 64) reduction(|| : miss)
    for (int t = 0; t < P.ntiles; ++t) {
        const int* un = unodes.data() + tile_uoff[t];
        for (int64_t s = tile_toff[t]; s < tile_toff[t + 1]; ++s) {
            const int a = tpairs[s] >> 8, bb = tpairs[s] & 0xff;
            const int64_t blk = block_of(un[a], un[bb]);
            if (blk < 0) miss = true;
            tblk[s] = blk;
        }
    }
    if (miss) throw Error(CVM_E_PATTERN_MISS,
        "Assembly: Gauss-point neighbour missing from the sparsity pattern; CSR mapping is inconsistent.");
    for (size_t s = 0; s < tblk.size(); ++s) kc_cnt[tblk[s] + 1]++;
    for (int64_t b = 0; b < P.nblocks; ++b) kc_cnt[b + 1] += kc_cnt[b];
    // block-major scratch: the tile kernel writes touched slot s straight to row
    // slot_dst[s], the position of its contribution in its block's segment; the
    // segments are in ascending tile order, i.e. GP order
    std::vector<uint32_t> slot_dst(tpairs.size());
    {
        std::vector<int64_t> fillp(kc_cnt.begin(), kc_cnt.end() - 1);
        for (size_t s = 0; s < tpairs.size(); ++s) slot_dst[s] = static_cast<uint32_t>(fillp[tblk[s]]++);
    }

    // F contributors per node
    std::vector<int> fc_cnt(nn + 1, 0);
    for (int v : unodes) fc_cnt[v + 1]++;
    for (int i = 0; i < nn; ++i) fc_cnt[i + 1] += fc_cnt[i];
    std::vector<int> fc_idx(fc_cnt[nn]);
    {
        std::vector<int> fillp(fc_cnt.begin(), fc_cnt.end() - 1);
        for (size_t k = 0; k < unodes.size(); ++k) fc_idx[fillp[unodes[k]]++] = static_cast<int>(k);
    }

    // fold contributors
    std::vector<int> fgp_fold, fgp_local, fgp_noff{0}, fgp_nodes, fgp_lpos, fgp_rpos;
    std::vector<int64_t> kf_cnt(P.nblocks + 1, 0);
    std::vector<int> ff_cnt(nn + 1, 0);
    for (const FGP& q : fgps) {
        fgp_fold.push_back(q.fold); fgp_local.push_back(q.local);
        const int base = fgp_noff.back();
        for (int v : q.nodes) { fgp_nodes.push_back(v); ff_cnt[v + 1]++; }
        fgp_noff.push_back(base + static_cast<int>(q.nodes.size()));
        const Fold& fo = c.folds[q.fold];
        Domain& L = *c.doms[fo.left];
        Domain& R = *c.doms[fo.right];
        for (int k = 0; k < q.left_n; ++k)
            fgp_lpos.push_back(static_cast<int>(std::lower_bound(q.nodes.begin(), q.nodes.end(), L.h_ids[q.left_b + k]) - q.nodes.begin()));
        for (int k = 0; k < q.right_n; ++k)
            fgp_rpos.push_back(static_cast<int>(std::lower_bound(q.nodes.begin(), q.nodes.end(), R.h_ids[q.right_b + k]) - q.nodes.begin()));
        for (size_t a = 0; a < q.nodes.size(); ++a)
            for (size_t bb = a; bb < q.nodes.size(); ++bb) kf_cnt[block_of(q.nodes[a], q.nodes[bb]) + 1]++;
    }
    for (int64_t b = 0; b < P.nblocks; ++b) kf_cnt[b + 1] += kf_cnt[b];
    std::vector<int> kf_idx(2 * kf_cnt[P.nblocks]);   // pairs: (global nodeset index lo, hi)
    {
        std::vector<int64_t> fillp(kf_cnt.begin(), kf_cnt.end() - 1);
        for (size_t qi = 0; qi < fgps.size(); ++qi) {
            const FGP& q = fgps[qi];
            const int base = fgp_noff[qi];
            for (size_t a = 0; a < q.nodes.size(); ++a)
                for (size_t bb = a; bb < q.nodes.size(); ++bb) {
                    const int64_t blk = block_of(q.nodes[a], q.nodes[bb]);
                    const int64_t at = fillp[blk]++;
                    kf_idx[2 * at] = base + static_cast<int>(a);
                    kf_idx[2 * at + 1] = base + static_cast<int>(bb);
                }
        }
    }
    for (int i = 0; i < nn; ++i) ff_cnt[i + 1] += ff_cnt[i];
    std::vector<int> ff_idx(ff_cnt[nn]);
    {
        std::vector<int> fillp(ff_cnt.begin(), ff_cnt.end() - 1);
        for (size_t k = 0; k < fgp_nodes.size(); ++k) ff_idx[fillp[fgp_nodes[k]]++] = static_cast<int>(k);
    }
    // nodeset entry -> crease GP
    std::vector<int> nd_gp(fgp_nodes.size());
    for (size_t qi = 0; qi < fgps.size(); ++qi)
        for (int k = fgp_noff[qi]; k < fgp_noff[qi + 1]; ++k) nd_gp[k] = static_cast<int>(qi);

    // ---- upload -----------------------------------------------------------------
    hipStream_t s = c.stream;
    P.d_blk_base_u.upload(base_u, s);
    P.d_blk_base_f.upload(base_f, s);
    P.d_blk_stride_u.upload(stride_u, s);
    P.d_blk_stride_f.upload(stride_f, s);
    P.d_tile_gp.upload(tile_gp, s);
    P.d_tile_uoff.upload(tile_uoff, s);
    P.d_tile_unodes.upload(unodes, s);
    P.d_tile_toff.upload(tile_toff, s);
    P.d_tile_tpairs.upload(tpairs, s);
    P.d_tile_nd.upload(tile_nd, s);
    P.d_pair_uloc.upload(uloc, s);
    P.d_pair_urank.upload(urank, s);
    P.d_kc_off.upload(kc_cnt, s);
    P.d_slot_dst.upload(slot_dst, s);
    P.d_fc_off.upload(fc_cnt, s);
    P.d_fc_idx.upload(fc_idx, s);
    if (P.nfgp) {
        P.d_fgp_fold.upload(fgp_fold, s);
        P.d_fgp_local.upload(fgp_local, s);
        P.d_fgp_noff.upload(fgp_noff, s);
        P.d_fgp_nodes.upload(fgp_nodes, s);
        P.d_fgp_lpos.upload(fgp_lpos, s);
        P.d_fgp_rpos.upload(fgp_rpos, s);
        P.d_fnd_gp.upload(nd_gp, s);
    }
    P.d_kf_off.upload(kf_cnt, s);
    P.d_kf_idx.upload(kf_idx.empty() ? std::vector<int>{0} : kf_idx, s);
    P.d_ff_off.upload(ff_cnt, s);
    P.d_ff_idx.upload(ff_idx.empty() ? std::vector<int>{0} : ff_idx, s);
    if (!fgps.empty()) {
        std::vector<int> lpo{0}, rpo{0};
        for (const FGP& q : fgps) { lpo.push_back(lpo.back() + q.left_n); rpo.push_back(rpo.back() + q.right_n); }
        P.d_fgp_lpo.upload(lpo, s);
        P.d_fgp_rpo.upload(rpo, s);
    }
    P.nfnodes = static_cast<int64_t>(fgp_nodes.size());
    P.face_ng = ng;
    P.face_npairs = fd.npairs;
    CVM_HIP(hipStreamSynchronize(s));
    upload_fold_descriptors(c);
    P.ready = true;
}

}  // namespace cvm
