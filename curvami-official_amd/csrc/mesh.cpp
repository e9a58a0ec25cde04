// Mesh input: Gmsh 4.1 ASCII reader, Gauss rules on its elements, crease
// polylines from its line elements and the default support radius.
//
// Host code (no device work): these produce the arrays the device path takes
// (nodes, integration domains, creases, rho). They follow
//   GmshReader::readMesh            core/include/curvami/GmshReader.h:32-146
//   Quadrature::generate            core/include/curvami/Quadrature.h:14-233
//   MeshUtils::extractCurve         core/include/curvami/MeshUtils.h:16-134
//   MeshUtils::computeSupportRadius core/include/curvami/MeshUtils.h:144-209
// Coordinates, weights and distances are formed with the reference's operand
// order and no FMA contraction (this file is built with -ffp-contract=off), so
// Gauss points, weights, tangents and rho come out bit-identical.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/curvami_gpu.h"

namespace {

thread_local std::string g_mesh_err;

struct MeshError : std::runtime_error {
    int code;
    MeshError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace

struct cvm_mesh {
    std::vector<double> X, Y;        // raw nodes in file order
    std::vector<int> type, entity;   // per element: Gmsh type and entity tag
    std::vector<int64_t> off;        // per element: first entry in idx (size nel + 1)
    std::vector<int> idx;            // element node indices (0-based, file order)
};

namespace {

// ---------------------------------------------------------------------------
// reader
// ---------------------------------------------------------------------------

// Whitespace tokenizer over the file text with line tracking for messages.
struct Lexer {
    const char* p;
    const char* end;
    const std::string& file;
    Lexer(const std::string& text, const std::string& f) : p(text.data()), end(text.data() + text.size()), file(f) {}

    void skip_ws() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p; }
    bool at_end() { skip_ws(); return p >= end; }
    // rest of the current line (trimmed); advances past the newline
    std::string line() {
        const char* s = p;
        while (p < end && *p != '\n') ++p;
        const char* e = p;
        if (p < end) ++p;
        while (s < e && (*s == ' ' || *s == '\t')) ++s;
        while (e > s && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\r')) --e;
        return std::string(s, e);
    }
    [[noreturn]] void fail(const char* what) {
        throw MeshError(CVM_E_INVALID, file + ": malformed mesh (" + what + ")");
    }
    long long integer(const char* what) {
        skip_ws();
        if (p >= end) fail(what);
        char* q = nullptr;
        errno = 0;
        const long long v = std::strtoll(p, &q, 10);
        if (q == p || errno != 0 || (q < end && !(*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n'))) fail(what);
        p = q;
        return v;
    }
    double real(const char* what) {
        skip_ws();
        if (p >= end) fail(what);
        char* q = nullptr;
        const double v = std::strtod(p, &q);   // correctly rounded, as istream >> double
        if (q == p || (q < end && !(*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n'))) fail(what);
        p = q;
        return v;
    }
};

int nodes_of_type(int t) {   // GmshReader.h:20-30
    switch (t) {
        case 15: return 1;   // point
        case 1: return 2;    // 2-node line
        case 2: return 3;    // 3-node triangle
        case 3: return 4;    // 4-node quadrangle
        default: return 0;
    }
}

void read_gmsh(const std::string& path, cvm_mesh& m) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw MeshError(CVM_E_INVALID, "could not open Gmsh file: " + path);
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);

    Lexer L(text, path);
    std::unordered_map<long long, int> tag_to_index;
    bool have_format = false;
    while (!L.at_end()) {
        const std::string ln = L.line();
        if (ln == "$MeshFormat") {
            const double version = L.real("format version");
            const long long ftype = L.integer("file type");
            L.integer("data size");
            if (!(version >= 4.0 && version < 5.0)) throw MeshError(CVM_E_INVALID, path + ": only Gmsh 4.x meshes are read");
            if (ftype != 0) throw MeshError(CVM_E_INVALID, path + ": binary Gmsh files are not read (save as ASCII msh41)");
            have_format = true;
        } else if (ln == "$Nodes") {   // GmshReader.h:47-82
            const long long nblocks = L.integer("node header");
            const long long nnodes = L.integer("node header");
            L.integer("node header"); L.integer("node header");
            if (nblocks < 0 || nnodes < 0) L.fail("node header");
            m.X.reserve(m.X.size() + nnodes);
            m.Y.reserve(m.Y.size() + nnodes);
            std::vector<long long> tags;
            for (long long b = 0; b < nblocks; ++b) {
                const long long dim = L.integer("node block");
                L.integer("node block");
                const long long parametric = L.integer("node block");
                const long long nb = L.integer("node block");
                if (nb < 0 || dim < 0 || dim > 3) L.fail("node block");
                tags.resize(static_cast<size_t>(nb));
                for (long long i = 0; i < nb; ++i) tags[i] = L.integer("node tag");
                for (long long i = 0; i < nb; ++i) {
                    const double x = L.real("node coordinate");
                    const double y = L.real("node coordinate");
                    L.real("node coordinate");                       // z: the sheet lies in z = 0
                    if (parametric) for (long long k = 0; k < dim; ++k) L.real("parametric coordinate");
                    tag_to_index[tags[i]] = static_cast<int>(m.X.size());
                    m.X.push_back(x);
                    m.Y.push_back(y);
                }
            }
        } else if (ln == "$Elements") {   // GmshReader.h:87-134
            const long long nblocks = L.integer("element header");
            const long long nel = L.integer("element header");
            L.integer("element header"); L.integer("element header");
            if (nblocks < 0 || nel < 0) L.fail("element header");
            if (m.off.empty()) m.off.push_back(0);
            for (long long b = 0; b < nblocks; ++b) {
                L.integer("element block");
                const long long etag = L.integer("element block");
                const long long etype = L.integer("element block");
                const long long nb = L.integer("element block");
                const int nv = nodes_of_type(static_cast<int>(etype));
                if (nv == 0)
                    throw MeshError(CVM_E_INVALID, path + ": unsupported Gmsh element type " + std::to_string(etype) +
                                                       "; cannot determine its node count");
                if (nb < 0) L.fail("element block");
                for (long long i = 0; i < nb; ++i) {
                    L.integer("element tag");
                    for (int j = 0; j < nv; ++j) {
                        const long long t = L.integer("element node tag");
                        auto it = tag_to_index.find(t);
                        if (it == tag_to_index.end())
                            throw MeshError(CVM_E_INVALID, path + ": element references undefined node tag " +
                                                               std::to_string(t));
                        m.idx.push_back(it->second);
                    }
                    m.type.push_back(static_cast<int>(etype));
                    m.entity.push_back(static_cast<int>(etag));
                    m.off.push_back(static_cast<int64_t>(m.idx.size()));
                }
            }
        }
    }
    if (!have_format) throw MeshError(CVM_E_INVALID, path + ": no $MeshFormat section (not a Gmsh file)");
    if (m.X.empty() || m.type.empty())
        throw MeshError(CVM_E_INVALID, path + " yielded " + std::to_string(m.X.size()) + " nodes and " +
                                           std::to_string(m.type.size()) + " elements; the file is empty or malformed");
    if (m.off.empty()) m.off.push_back(0);
}

// ---------------------------------------------------------------------------
// quadrature (Quadrature.h:14-233)
// ---------------------------------------------------------------------------

struct Rules {
    std::vector<double> lp, lw, qp, qw;
    struct Tri { double u, v, w, wt; };
    std::vector<Tri> tri;
    Rules(int line_degree, int face_degree) {
        if (line_degree <= 1) { lp = {0.0}; lw = {2.0}; }                      // :18-20
        else {                                                                  // 3-point Gauss-Legendre
            lp = {-0.77459666924148337704, 0.0, 0.77459666924148337704};
            lw = {0.55555555555555555556, 0.88888888888888888889, 0.55555555555555555556};
        }
        if (face_degree <= 1) { qp = {0.0}; qw = {2.0}; }                      // :33-43
        else if (face_degree == 2) { qp = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)}; qw = {1.0, 1.0}; }
        else {
            qp = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)};
            qw = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
        }
        if (face_degree <= 1) tri = {{1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0}}; // :46-57
        else tri = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0},
                    {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 3.0},
                    {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0}};
    }
};

struct QuadOut {
    std::vector<int> active;                          // raw indices of face-element nodes, ascending
    std::vector<double> face, edge, edge_tn, point;   // x, y, w per GP; tx, ty, nx, ny per edge GP
    std::vector<int> face_tag, edge_tag, point_tag;
};

void quadrature(const cvm_mesh& m, int line_degree, int face_degree, QuadOut& q) {
    const Rules R(line_degree, face_degree);
    std::set<int> active;
    const size_t nel = m.type.size();
    for (size_t e = 0; e < nel; ++e) {
        const int* v = m.idx.data() + m.off[e];
        const int tag = m.entity[e];
        switch (m.type[e]) {
        case 15: {                                                     // :87-97
            q.point.insert(q.point.end(), {m.X[v[0]], m.Y[v[0]], 1.0});
            q.point_tag.push_back(tag);
            break;
        }
        case 1: {                                                      // :100-128
            const double p1x = m.X[v[0]], p1y = m.Y[v[0]], p2x = m.X[v[1]], p2y = m.Y[v[1]];
            double tx = p2x - p1x, ty = p2y - p1y;
            const double L = std::sqrt(tx * tx + ty * ty);
            if (!(L > 0.0)) break;
            tx /= L; ty /= L;
            const double nx = -ty, ny = tx;
            const double detJ = L / 2.0;
            for (size_t i = 0; i < R.lp.size(); ++i) {
                const double xi = R.lp[i];
                const double a = (1.0 - xi) / 2.0, b = (1.0 + xi) / 2.0;
                q.edge.insert(q.edge.end(), {a * p1x + b * p2x, a * p1y + b * p2y, detJ * R.lw[i]});
                q.edge_tn.insert(q.edge_tn.end(), {tx, ty, nx, ny});
                q.edge_tag.push_back(tag);
            }
            break;
        }
        case 2: {                                                      // :131-152
            const double x1 = m.X[v[0]], y1 = m.Y[v[0]], x2 = m.X[v[1]], y2 = m.Y[v[1]];
            const double x3 = m.X[v[2]], y3 = m.Y[v[2]];
            const double area = 0.5 * std::abs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
            for (const auto& t : R.tri) {
                q.face.insert(q.face.end(), {t.u * x1 + t.v * x2 + t.w * x3, t.u * y1 + t.v * y2 + t.w * y3,
                                             area * t.wt});
                q.face_tag.push_back(tag);
            }
            for (int k = 0; k < 3; ++k) active.insert(v[k]);
            break;
        }
        case 3: {                                                      // :155-217
            double px[4], py[4];
            for (int k = 0; k < 4; ++k) { px[k] = m.X[v[k]]; py[k] = m.Y[v[k]]; active.insert(v[k]); }
            for (size_t ix = 0; ix < R.qp.size(); ++ix) {
                const double xi = R.qp[ix], wx = R.qw[ix];
                for (size_t iy = 0; iy < R.qp.size(); ++iy) {
                    const double eta = R.qp[iy], wy = R.qw[iy];
                    const double N1 = 0.25 * (1.0 - xi) * (1.0 - eta), N2 = 0.25 * (1.0 + xi) * (1.0 - eta);
                    const double N3 = 0.25 * (1.0 + xi) * (1.0 + eta), N4 = 0.25 * (1.0 - xi) * (1.0 + eta);
                    // bilinear map Jacobian, rows d/dxi and d/deta
                    const double J11 = -0.25 * (1.0 - eta) * px[0] + 0.25 * (1.0 - eta) * px[1] +
                                       0.25 * (1.0 + eta) * px[2] - 0.25 * (1.0 + eta) * px[3];
                    const double J12 = -0.25 * (1.0 - eta) * py[0] + 0.25 * (1.0 - eta) * py[1] +
                                       0.25 * (1.0 + eta) * py[2] - 0.25 * (1.0 + eta) * py[3];
                    const double J21 = -0.25 * (1.0 - xi) * px[0] - 0.25 * (1.0 + xi) * px[1] +
                                       0.25 * (1.0 + xi) * px[2] + 0.25 * (1.0 - xi) * px[3];
                    const double J22 = -0.25 * (1.0 - xi) * py[0] - 0.25 * (1.0 + xi) * py[1] +
                                       0.25 * (1.0 + xi) * py[2] + 0.25 * (1.0 - xi) * py[3];
                    const double detJ = std::abs(J11 * J22 - J12 * J21);   // clockwise quads: |detJ|
                    q.face.insert(q.face.end(), {N1 * px[0] + N2 * px[1] + N3 * px[2] + N4 * px[3],
                                                 N1 * py[0] + N2 * py[1] + N3 * py[2] + N4 * py[3], detJ * wx * wy});
                    q.face_tag.push_back(tag);
                }
            }
            break;
        }
        default: break;
        }
    }
    q.active.assign(active.begin(), active.end());
}

// ---------------------------------------------------------------------------
// crease polylines (MeshUtils.h:16-134)
// ---------------------------------------------------------------------------

double cross(double ox, double oy, double ax, double ay, double bx, double by) {
    return (ax - ox) * (by - oy) - (ay - oy) * (bx - ox);
}

// Andrew's monotone chain, collinear points dropped (MeshUtils.h:22-39)
std::vector<std::pair<double, double>> convex_hull(std::vector<std::pair<double, double>> P) {
    const size_t n = P.size();
    if (n <= 3) return P;
    std::sort(P.begin(), P.end());   // x, then y
    std::vector<std::pair<double, double>> H(2 * n);
    size_t k = 0;
    auto turn = [&](size_t a, size_t b, const std::pair<double, double>& c) {
        return cross(H[a].first, H[a].second, H[b].first, H[b].second, c.first, c.second);
    };
    for (size_t i = 0; i < n; ++i) {
        while (k >= 2 && turn(k - 2, k - 1, P[i]) <= 0) --k;
        H[k++] = P[i];
    }
    for (size_t i = n - 1, t = k + 1; i > 0; --i) {
        while (k >= t && turn(k - 2, k - 1, P[i - 1]) <= 0) --k;
        H[k++] = P[i - 1];
    }
    H.resize(k - 1);
    return H;
}

bool on_hull(double x, double y, const std::vector<std::pair<double, double>>& H, double tol) {   // :42-55
    for (size_t i = 0; i < H.size(); ++i) {
        const double ax = H[i].first, ay = H[i].second;
        const double bx = H[(i + 1) % H.size()].first, by = H[(i + 1) % H.size()].second;
        const double abx = bx - ax, aby = by - ay;
        const double len2 = abx * abx + aby * aby;
        if (len2 == 0) continue;
        const double t = std::max(0.0, std::min(1.0, ((x - ax) * abx + (y - ay) * aby) / len2));
        const double dx = x - (ax + t * abx), dy = y - (ay + t * aby);
        if (std::sqrt(dx * dx + dy * dy) <= tol) return true;
    }
    return false;
}

struct Curve {
    std::vector<int> nodes;   // walk order (a closed chain repeats its first node at the end)
    bool closed = false, ext_s = false, ext_e = false;
    int64_t dropped = 0;
};

Curve extract_curve(const cvm_mesh& m, const int* tags, int ntags) {
    Curve c;
    std::map<int, std::vector<int>> adj;   // ordered by node index: the start node is the lowest end
    const size_t nel = m.type.size();
    for (size_t e = 0; e < nel; ++e) {
        if (m.type[e] != 1 || std::find(tags, tags + ntags, m.entity[e]) == tags + ntags) continue;
        const int a = m.idx[m.off[e]], b = m.idx[m.off[e] + 1];
        adj[a].push_back(b);
        adj[b].push_back(a);
    }
    if (adj.empty()) return c;
    int start = -1;
    for (const auto& kv : adj) if (kv.second.size() == 1) { start = kv.first; break; }
    if (start < 0) start = adj.begin()->first;
    // walk: first neighbour that is not the one we came from. A branch point can
    // lead the walk into a cycle that does not pass the start; the visited set
    // ends it there (the reference's walk would not terminate).
    std::set<int> seen;
    int cur = start, prev = -1;
    while (true) {
        c.nodes.push_back(cur);
        seen.insert(cur);
        int next = -1;
        for (int nb : adj[cur]) if (nb != prev) { next = nb; break; }
        if (next == -1 || next == start) {
            if (next == start) c.closed = true;
            break;
        }
        if (seen.count(next)) break;
        prev = cur;
        cur = next;
    }
    c.dropped = static_cast<int64_t>(adj.size()) - static_cast<int64_t>(c.nodes.size());
    std::vector<std::pair<double, double>> pts(m.X.size());
    for (size_t i = 0; i < m.X.size(); ++i) pts[i] = {m.X[i], m.Y[i]};
    const auto H = convex_hull(pts);
    const int first = c.nodes.front(), last = c.nodes.back();
    c.ext_s = on_hull(m.X[first], m.Y[first], H, 1e-4);
    c.ext_e = on_hull(m.X[last], m.Y[last], H, 1e-4);
    if (c.closed) c.nodes.push_back(c.nodes.front());
    return c;
}

// ---------------------------------------------------------------------------
// support radius (MeshUtils.h:144-209): (order + 3) x median over nodes of the
// distance to the 3rd-nearest other node. Bucket grid + growing square rings;
// a ring r beyond the cell holds nothing closer than r * cell, which bounds
// the search. The result is one of the pairwise distances sqrt(dx^2 + dy^2),
// so any complete search returns the same bits.
// ---------------------------------------------------------------------------

double support_radius(const double* X, const double* Y, int64_t N, int order) {
    if (N < 4) return 0.1;
    double lox = X[0], loy = Y[0], hix = X[0], hiy = Y[0];
    for (int64_t i = 0; i < N; ++i) {
        lox = std::min(lox, X[i]); hix = std::max(hix, X[i]);
        loy = std::min(loy, Y[i]); hiy = std::max(hiy, Y[i]);
    }
    const double ex = std::max(hix - lox, 1e-30), ey = std::max(hiy - loy, 1e-30);
    double cell = std::sqrt(ex * ey / static_cast<double>(N)) * 1.5;
    if (!(cell > 0.0)) cell = std::max(ex, ey);
    const int nx = std::max(1, static_cast<int>(std::min(ex / cell, 1e6)) + 1);
    const int ny = std::max(1, static_cast<int>(std::min(ey / cell, 1e6)) + 1);
    auto cx_of = [&](double x) { return std::min(nx - 1, std::max(0, static_cast<int>((x - lox) / cell))); };
    auto cy_of = [&](double y) { return std::min(ny - 1, std::max(0, static_cast<int>((y - loy) / cell))); };
    // counting sort of the nodes into cells
    const size_t ncell = static_cast<size_t>(nx) * ny;
    std::vector<int64_t> start(ncell + 1, 0);
    std::vector<int64_t> order_idx(static_cast<size_t>(N));
    for (int64_t i = 0; i < N; ++i) ++start[static_cast<size_t>(cy_of(Y[i])) * nx + cx_of(X[i]) + 1];
    for (size_t k = 0; k < ncell; ++k) start[k + 1] += start[k];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < N; ++i) order_idx[fill[static_cast<size_t>(cy_of(Y[i])) * nx + cx_of(X[i])]++] = i;
    }
    std::vector<double> h(static_cast<size_t>(N));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const double px = X[i], py = Y[i];
        const int cx = cx_of(px), cy = cy_of(py);
        const double inf = std::numeric_limits<double>::infinity();
        double best[4] = {inf, inf, inf, inf};   // ascending; the node itself gives 0
        auto scan = [&](int gx, int gy) {
            if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) return;
            const size_t k = static_cast<size_t>(gy) * nx + gx;
            for (int64_t s = start[k]; s < start[k + 1]; ++s) {
                const int64_t j = order_idx[s];
                const double dx = px - X[j], dy = py - Y[j];
                const double d = std::sqrt(dx * dx + dy * dy);
                if (!(d < best[3])) continue;
                int q = 3;
                while (q > 0 && best[q - 1] > d) { best[q] = best[q - 1]; --q; }
                best[q] = d;
            }
        };
        for (int r = 0;; ++r) {
            const int x0 = cx - r, x1 = cx + r, y0 = cy - r, y1 = cy + r;
            if (r == 0) {
                scan(cx, cy);
            } else {   // the ring of cells at Chebyshev distance r
                for (int gx = x0; gx <= x1; ++gx) { scan(gx, y0); scan(gx, y1); }
                for (int gy = y0 + 1; gy < y1; ++gy) { scan(x0, gy); scan(x1, gy); }
            }
            if (best[3] <= static_cast<double>(r) * cell) break;
            if (x0 < 0 && y0 < 0 && x1 >= nx && y1 >= ny) break;   // every cell seen
        }
        h[static_cast<size_t>(i)] = best[3];
    }
    std::nth_element(h.begin(), h.begin() + N / 2, h.end());
    return static_cast<double>(order + 3) * h[static_cast<size_t>(N / 2)];
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return CVM_OK;
    } catch (const MeshError& e) {
        g_mesh_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_mesh_err = "out of host memory";
        return CVM_E_NO_MEMORY;
    } catch (const std::exception& e) {
        g_mesh_err = e.what();
        return CVM_E_INVALID;
    }
}

}  // namespace

extern "C" {

const char* cvm_mesh_last_error(void) { return g_mesh_err.c_str(); }

int cvm_mesh_read(const char* path, cvm_mesh** out) {
    return guarded([&] {
        if (!path || !out) throw MeshError(CVM_E_INVALID, "null argument");
        *out = nullptr;
        auto* m = new cvm_mesh();
        try {
            read_gmsh(path, *m);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

void cvm_mesh_destroy(cvm_mesh* m) { delete m; }

int cvm_mesh_info(const cvm_mesh* m, int64_t* out3) {
    return guarded([&] {
        if (!m || !out3) throw MeshError(CVM_E_INVALID, "null argument");
        out3[0] = static_cast<int64_t>(m->X.size());
        out3[1] = static_cast<int64_t>(m->type.size());
        out3[2] = static_cast<int64_t>(m->idx.size());
    });
}

int cvm_mesh_nodes(const cvm_mesh* m, double* X, double* Y) {
    return guarded([&] {
        if (!m) throw MeshError(CVM_E_INVALID, "null mesh");
        if (X) std::copy(m->X.begin(), m->X.end(), X);
        if (Y) std::copy(m->Y.begin(), m->Y.end(), Y);
    });
}

int cvm_mesh_elements(const cvm_mesh* m, int* type, int* entity_tag, int64_t* node_off, int* node_idx) {
    return guarded([&] {
        if (!m) throw MeshError(CVM_E_INVALID, "null mesh");
        if (type) std::copy(m->type.begin(), m->type.end(), type);
        if (entity_tag) std::copy(m->entity.begin(), m->entity.end(), entity_tag);
        if (node_off) std::copy(m->off.begin(), m->off.end(), node_off);
        if (node_idx) std::copy(m->idx.begin(), m->idx.end(), node_idx);
    });
}

int cvm_mesh_quadrature_sizes(const cvm_mesh* m, int line_degree, int face_degree, int64_t* out4) {
    return guarded([&] {
        if (!m || !out4) throw MeshError(CVM_E_INVALID, "null argument");
        QuadOut q;
        quadrature(*m, line_degree, face_degree, q);
        out4[0] = static_cast<int64_t>(q.active.size());
        out4[1] = static_cast<int64_t>(q.face_tag.size());
        out4[2] = static_cast<int64_t>(q.edge_tag.size());
        out4[3] = static_cast<int64_t>(q.point_tag.size());
    });
}

int cvm_mesh_quadrature(const cvm_mesh* m, int line_degree, int face_degree, int* active, double* face_xyw,
                        int* face_tag, double* edge_xyw, double* edge_tn, int* edge_tag, double* point_xyw,
                        int* point_tag) {
    return guarded([&] {
        if (!m) throw MeshError(CVM_E_INVALID, "null mesh");
        QuadOut q;
        quadrature(*m, line_degree, face_degree, q);
        auto put = [](const auto& v, auto* dst) { if (dst) std::copy(v.begin(), v.end(), dst); };
        put(q.active, active);
        put(q.face, face_xyw); put(q.face_tag, face_tag);
        put(q.edge, edge_xyw); put(q.edge_tn, edge_tn); put(q.edge_tag, edge_tag);
        put(q.point, point_xyw); put(q.point_tag, point_tag);
    });
}

int cvm_mesh_extract_curve(const cvm_mesh* m, const int* tags, int ntags, int64_t cap, double* vx, double* vy,
                           int64_t* nvert, int* flags, int64_t* dropped) {
    return guarded([&] {
        if (!m || (!tags && ntags > 0) || ntags < 0 || !nvert) throw MeshError(CVM_E_INVALID, "null argument");
        const Curve c = extract_curve(*m, tags, ntags);
        *nvert = static_cast<int64_t>(c.nodes.size());
        if (flags) *flags = (c.closed ? 1 : 0) | (c.ext_s ? 2 : 0) | (c.ext_e ? 4 : 0);
        if (dropped) *dropped = c.dropped;
        if (!vx && !vy) return;
        if (cap < *nvert) throw MeshError(CVM_E_INVALID, "vertex buffer too small");
        for (size_t k = 0; k < c.nodes.size(); ++k) {
            if (vx) vx[k] = m->X[c.nodes[k]];
            if (vy) vy[k] = m->Y[c.nodes[k]];
        }
    });
}

int cvm_support_radius(const double* X, const double* Y, int64_t n, int order, double* rho) {
    return guarded([&] {
        if (!rho || n < 0 || (n > 0 && (!X || !Y))) throw MeshError(CVM_E_INVALID, "null argument");
        if (order < 0) throw MeshError(CVM_E_INVALID, "order must be >= 0");
        *rho = support_radius(X, Y, n, order);
    });
}

}  // extern "C"
