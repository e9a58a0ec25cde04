"""CPU restatement of the reference's mesh input. TEST INFRASTRUCTURE ONLY:
imported by tests/ as the checker of the library's mesh entry points; the
product library never uses it.

Follows, function by function:
  read_gmsh       GmshReader::readMesh       core/include/curvami/GmshReader.h:32-146
  quadrature      Quadrature::generate       core/include/curvami/Quadrature.h:14-233
  convex_hull     MeshUtils::computeConvexHull core/include/curvami/MeshUtils.h:22-39
  on_hull         MeshUtils::isOnHullBoundary  core/include/curvami/MeshUtils.h:42-55
  extract_curve   MeshUtils::extractCurve    core/include/curvami/MeshUtils.h:57-134
  support_radius  MeshUtils::computeSupportRadius core/include/curvami/MeshUtils.h:144-209

Python floats are IEEE doubles evaluated left to right without contraction,
so the coordinates and weights below follow the reference's operand order and
are comparable bit for bit. support_radius is a brute-force all-pairs search
(a different algorithm from the reference's bucket grid with the same answer).

Pins (tests/test_mesh_input.py): node/element counts and coordinates against
the mesh files and the geometry their generating scripts state; Gauss rules
against closed-form integrals of monomials (quad 3x3: degree 5 per variable;
triangle 3-point: degree 2; line 3-point: degree 5); weight sums against the
domain areas and crease lengths; the hull against scipy.spatial.ConvexHull;
the support radius against the closed form (order + 3) h of a uniform grid.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

NODES_PER_TYPE = {15: 1, 1: 2, 2: 3, 3: 4}   # GmshReader.h:20-30


class GmshMesh:
    def __init__(self):
        self.X: List[float] = []
        self.Y: List[float] = []
        self.elements: List[Tuple[int, int, List[int]]] = []   # (type, entity tag, node indices)


def read_gmsh(path: str) -> GmshMesh:
    """GmshReader.h:32-146 on a whitespace token stream."""
    with open(path, "r") as f:
        lines = f.read().split("\n")
    m = GmshMesh()
    tag_to_index: Dict[int, int] = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if ln not in ("$Nodes", "$Elements"):
            continue
        end_marker = "$EndNodes" if ln == "$Nodes" else "$EndElements"
        toks: List[str] = []
        while lines[i].strip() != end_marker:
            toks.extend(lines[i].split())
            i += 1
        it = iter(toks)
        nxt = lambda: next(it)
        if ln == "$Nodes":
            nblocks = int(nxt()); int(nxt()); int(nxt()); int(nxt())
            for _ in range(nblocks):
                dim = int(nxt()); int(nxt()); param = int(nxt()); nb = int(nxt())
                tags = [int(nxt()) for _ in range(nb)]
                for t in tags:
                    x = float(nxt()); y = float(nxt()); float(nxt())
                    if param:
                        for _ in range(dim):
                            float(nxt())
                    tag_to_index[t] = len(m.X)
                    m.X.append(x)
                    m.Y.append(y)
        else:
            nblocks = int(nxt()); int(nxt()); int(nxt()); int(nxt())
            for _ in range(nblocks):
                int(nxt()); etag = int(nxt()); etype = int(nxt()); nb = int(nxt())
                nv = NODES_PER_TYPE.get(etype, 0)
                if nv == 0:
                    raise ValueError(f"unsupported Gmsh element type {etype}")
                for _ in range(nb):
                    int(nxt())
                    idx = []
                    for _ in range(nv):
                        t = int(nxt())
                        if t not in tag_to_index:
                            raise ValueError(f"element references undefined node tag {t}")
                        idx.append(tag_to_index[t])
                    m.elements.append((etype, etag, idx))
    if not m.X or not m.elements:
        raise ValueError("empty or malformed mesh")
    return m


def _rules(line_degree: int, face_degree: int):
    """Quadrature.h:14-57."""
    if line_degree <= 1:
        lp, lw = [0.0], [2.0]
    else:
        lp = [-0.77459666924148337704, 0.0, 0.77459666924148337704]
        lw = [0.55555555555555555556, 0.88888888888888888889, 0.55555555555555555556]
    if face_degree <= 1:
        qp, qw = [0.0], [2.0]
    elif face_degree == 2:
        qp, qw = [-1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0)], [1.0, 1.0]
    else:
        qp, qw = [-math.sqrt(3.0 / 5.0), 0.0, math.sqrt(3.0 / 5.0)], [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]
    if face_degree <= 1:
        tri = [(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0)]
    else:
        tri = [(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0),
               (1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 3.0),
               (1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0)]
    return lp, lw, qp, qw, tri


def quadrature(m: GmshMesh, line_degree: int = 3, face_degree: int = 2) -> dict:
    """Quadrature.h:59-233, element order. Returns arrays: active (sorted raw
    indices of face-element nodes), face (x, y, w), face_tag, edge (x, y, w),
    edge_tn (tx, ty, nx, ny), edge_tag, point (x, y, w), point_tag."""
    lp, lw, qp, qw, tri = _rules(line_degree, face_degree)
    X, Y = m.X, m.Y
    face, face_tag, edge, edge_tn, edge_tag, point, point_tag = [], [], [], [], [], [], []
    active = set()
    for etype, tag, v in m.elements:
        if etype == 15:
            point.append((X[v[0]], Y[v[0]], 1.0)); point_tag.append(tag)
        elif etype == 1:
            x1, y1, x2, y2 = X[v[0]], Y[v[0]], X[v[1]], Y[v[1]]
            tx, ty = x2 - x1, y2 - y1
            L = math.sqrt(tx * tx + ty * ty)
            if not L > 0.0:
                continue
            tx, ty = tx / L, ty / L
            detJ = L / 2.0
            for xi, wt in zip(lp, lw):
                a, b = (1.0 - xi) / 2.0, (1.0 + xi) / 2.0
                edge.append((a * x1 + b * x2, a * y1 + b * y2, detJ * wt))
                edge_tn.append((tx, ty, -ty, tx)); edge_tag.append(tag)
        elif etype == 2:
            (x1, x2, x3), (y1, y2, y3) = [X[k] for k in v], [Y[k] for k in v]
            area = 0.5 * abs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2))
            for u, vv, w, wt in tri:
                face.append((u * x1 + vv * x2 + w * x3, u * y1 + vv * y2 + w * y3, area * wt))
                face_tag.append(tag)
            active.update(v)
        elif etype == 3:
            px, py = [X[k] for k in v], [Y[k] for k in v]
            active.update(v)
            for xi, wx in zip(qp, qw):
                for eta, wy in zip(qp, qw):
                    N = (0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                         0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta))
                    dxi = (-0.25 * (1.0 - eta), 0.25 * (1.0 - eta), 0.25 * (1.0 + eta), -0.25 * (1.0 + eta))
                    deta = (-0.25 * (1.0 - xi), -0.25 * (1.0 + xi), 0.25 * (1.0 + xi), 0.25 * (1.0 - xi))
                    J11 = dxi[0] * px[0] + dxi[1] * px[1] + dxi[2] * px[2] + dxi[3] * px[3]
                    J12 = dxi[0] * py[0] + dxi[1] * py[1] + dxi[2] * py[2] + dxi[3] * py[3]
                    J21 = deta[0] * px[0] + deta[1] * px[1] + deta[2] * px[2] + deta[3] * px[3]
                    J22 = deta[0] * py[0] + deta[1] * py[1] + deta[2] * py[2] + deta[3] * py[3]
                    detJ = abs(J11 * J22 - J12 * J21)
                    face.append((N[0] * px[0] + N[1] * px[1] + N[2] * px[2] + N[3] * px[3],
                                 N[0] * py[0] + N[1] * py[1] + N[2] * py[2] + N[3] * py[3], detJ * wx * wy))
                    face_tag.append(tag)
    a = lambda rows, k: np.array(rows, dtype=np.float64).reshape(-1, k)
    return {"active": np.array(sorted(active), dtype=np.int64),
            "face": a(face, 3), "face_tag": np.array(face_tag, np.int64),
            "edge": a(edge, 3), "edge_tn": a(edge_tn, 4), "edge_tag": np.array(edge_tag, np.int64),
            "point": a(point, 3), "point_tag": np.array(point_tag, np.int64)}


def _cross(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def convex_hull(pts):
    """MeshUtils.h:22-39: monotone chain, collinear points dropped,
    counter-clockwise from the lowest-x point."""
    P = sorted(pts)
    n = len(P)
    if n <= 3:
        return list(pts)
    lower, upper = [], []
    for p in P:
        while len(lower) >= 2 and _cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(P):
        while len(upper) >= 2 and _cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


def on_hull(p, hull, tol=1e-4) -> bool:
    """MeshUtils.h:42-55."""
    for i in range(len(hull)):
        A, B = hull[i], hull[(i + 1) % len(hull)]
        abx, aby = B[0] - A[0], B[1] - A[1]
        len2 = abx * abx + aby * aby
        if len2 == 0:
            continue
        t = max(0.0, min(1.0, ((p[0] - A[0]) * abx + (p[1] - A[1]) * aby) / len2))
        dx, dy = p[0] - (A[0] + t * abx), p[1] - (A[1] + t * aby)
        if math.sqrt(dx * dx + dy * dy) <= tol:
            return True
    return False


def extract_curve(m: GmshMesh, tags):
    """MeshUtils.h:57-134: (vx, vy, closed, extend_start, extend_end, dropped)."""
    adj: Dict[int, List[int]] = {}
    for etype, tag, v in m.elements:
        if etype == 1 and tag in tags:
            adj.setdefault(v[0], []).append(v[1])
            adj.setdefault(v[1], []).append(v[0])
    if not adj:
        return np.zeros(0), np.zeros(0), False, False, False, 0
    keys = sorted(adj)
    ends = [k for k in keys if len(adj[k]) == 1]
    start = ends[0] if ends else keys[0]
    order, seen, prev, cur, closed = [], set(), -1, start, False
    while True:
        order.append(cur); seen.add(cur)
        nxt = next((nb for nb in adj[cur] if nb != prev), -1)
        if nxt == -1 or nxt == start:
            closed = nxt == start
            break
        if nxt in seen:       # branch back into the chain: stop
            break
        prev, cur = cur, nxt
    dropped = len(adj) - len(order)
    hull = convex_hull(list(zip(m.X, m.Y)))
    ext_s = on_hull((m.X[order[0]], m.Y[order[0]]), hull)
    ext_e = on_hull((m.X[order[-1]], m.Y[order[-1]]), hull)
    if closed:
        order.append(order[0])
    return (np.array([m.X[k] for k in order]), np.array([m.Y[k] for k in order]),
            closed, ext_s, ext_e, dropped)


def support_radius(X, Y, order: int) -> float:
    """MeshUtils.h:144-209 by brute force: (order + 3) x the element n//2 of the
    ascending distances to each node's 3rd-nearest other node."""
    X = np.asarray(X, np.float64); Y = np.asarray(Y, np.float64)
    n = X.size
    if n < 4:
        return 0.1
    h = np.empty(n)
    for s in range(0, n, 512):
        dx = X[s:s + 512, None] - X[None, :]
        dy = Y[s:s + 512, None] - Y[None, :]
        d = np.sqrt(dx * dx + dy * dy)
        h[s:s + 512] = np.partition(d, 3, axis=1)[:, 3]   # index 0 is the node itself
    return float(order + 3) * float(np.partition(h, n // 2)[n // 2])
