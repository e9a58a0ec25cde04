"""Closed-form generators for creased-sheet inputs.

The reference reads Gmsh meshes and builds quadrature from them
(core/include/curvami/GmshReader.h:32-146, core/include/curvami/Quadrature.h:59-233);
here the same kinds of arrays are generated directly for square sheets on a
uniform node grid, so that the GPU path and the CPU oracle see identical,
seeded inputs:

* nodes on an n x n grid of the unit square, id = j*n + i (x fastest);
* face Gauss points from an m x m tensor Gauss-Legendre rule per background
  cell, cells in row-major order (cy outer, cx inner), points ix outer, iy
  inner, as in the quad loop of Quadrature.h:163-169;
* creases as polylines (Enrichment::FoldLine, core/include/curvami/Enrichment.h:10-15)
  with extend flags set when an end lies on the sheet boundary
  (MeshUtils::extractCurve, core/include/curvami/MeshUtils.h:120-130);
* crease Gauss points at segment midpoints with the 1-point line rule
  (weight = segment length, tangent along the segment, normal = tangent
  turned by +90 degrees; Quadrature.h:99-121), and the fold's left/right
  Gauss points offset by +/-1e-7 along the normal (plate/Meshless.h:30-46);
* face Gauss points that lie within 1e-9 of a crease split into two
  half-weight twins at +/-1e-7 (plate/SetupHelper.h:60-106);
* a single isotropic ply with G = E / (2 (1 + nu)) (main.cpp:54-61);
* displacements from SplitMix64 + Box-Muller (explicitly specified, so that
  every consumer gets the same numbers).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Tuple

import numpy as np

# Gauss-Legendre abscissae / weights on [-1, 1]
_GL = {
    1: ([0.0], [2.0]),
    2: ([-0.5773502691896257, 0.5773502691896257], [1.0, 1.0]),
    3: ([-0.7745966692414834, 0.0, 0.7745966692414834],
        [0.5555555555555556, 0.8888888888888888, 0.5555555555555556]),
    4: ([-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526],
        [0.3478548451374638, 0.6521451548625461, 0.6521451548625461, 0.3478548451374638]),
}

SPLIT_EPS = 1e-9     # plate/SetupHelper.h:62
OFFSET = 1e-7        # plate/Meshless.h:37, plate/SetupHelper.h:62


@dataclasses.dataclass
class Crease:
    vx: np.ndarray
    vy: np.ndarray
    ext_s: bool = True
    ext_e: bool = True
    closed: bool = False


@dataclasses.dataclass
class GPSet:
    gx: np.ndarray
    gy: np.ndarray
    w: np.ndarray
    nx: Optional[np.ndarray] = None
    ny: Optional[np.ndarray] = None
    tx: Optional[np.ndarray] = None
    ty: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.gx.shape[0])


@dataclasses.dataclass
class Fold:
    """plate/PlateData.h:73-94: hinge of stiffness Kt and rest angle Ang0 (fold
    sign applied) between the left / right crease GPs; rigid folds are held by
    angle rows; force_left / force_right are dead loads on those GPs."""
    left: GPSet
    right: GPSet
    Kt: float
    Ang0: float = 0.0
    crease: int = 0
    rigid: bool = False
    force_left: tuple = (0.0, 0.0, 0.0)
    force_right: tuple = (0.0, 0.0, 0.0)


@dataclasses.dataclass
class Boundary:
    """An edge or point integration domain with its BoundaryCondition
    (plate/PlateData.h:36-56). active: bit 0 u, 1 v, 2 w, 3 normal rotation."""
    gps: GPSet
    kind: str                 # "point" or "edge"
    active: int = 0
    val: tuple = (0.0, 0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    moment: float = 0.0


@dataclasses.dataclass
class Sheet:
    name: str
    n_side: int
    X: np.ndarray
    Y: np.ndarray
    face: GPSet
    creases: List[Crease]
    folds: List[Fold]
    rho: float
    order: int = 2
    kernel: int = 0          # 0 quartic spline (reference), 1 cubic B-spline
    E: float = 1.0
    nu: float = 0.3
    t: float = 0.01
    n_split: int = 0
    boundaries: List[Boundary] = dataclasses.field(default_factory=list)
    # natural curvature band (x0, width, theta0 [rad], sign: -1 mountain, +1 valley), or None
    natural_curvature: Optional[Tuple[float, float, float, float]] = None

    @property
    def nn(self) -> int:
        return int(self.X.shape[0])

    @property
    def plies(self) -> np.ndarray:
        """One isotropic ply: E1, E2, nu12, G12, angle_deg, thickness (main.cpp:58-60)."""
        G = self.E / (2.0 * (1.0 + self.nu))
        return np.array([[self.E, self.E, self.nu, G, 0.0, self.t]], dtype=np.float64)


# ---------------------------------------------------------------------------
# Random numbers: SplitMix64 + Box-Muller
# ---------------------------------------------------------------------------
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs for state `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Uniforms in (0, 1]: ((z >> 11) + 1) * 2^-53."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, sigma: float = 1.0) -> np.ndarray:
    """Box-Muller normals: pairs (u1, u2) -> r cos(2 pi u2), r sin(2 pi u2)."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return sigma * z[:n]


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------
def grid_nodes(n_side: int, L: float = 1.0):
    h = L / (n_side - 1)
    i = np.arange(n_side, dtype=np.float64)
    X = np.tile(i * h, n_side)
    Y = np.repeat(i * h, n_side)
    return X, Y, h


def face_gauss(n_side: int, h: float, q: int = 4) -> GPSet:
    pts, wts = _GL[q]
    pts = np.asarray(pts); wts = np.asarray(wts)
    nc = n_side - 1
    c = np.arange(nc, dtype=np.float64)
    cx = np.tile(c, nc)          # cells row-major: cy outer, cx inner
    cy = np.repeat(c, nc)
    x0 = cx * h; y0 = cy * h
    x1 = (cx + 1.0) * h; y1 = (cy + 1.0) * h
    # within a cell: ix outer, iy inner (Quadrature.h:163-169)
    xi = np.repeat(pts, q); eta = np.tile(pts, q)
    wx = np.repeat(wts, q); wy = np.tile(wts, q)
    # bilinear map of the cell (Quadrature.h:175-211): X = sum N_a p_a
    N1 = 0.25 * (1.0 - xi) * (1.0 - eta)
    N2 = 0.25 * (1.0 + xi) * (1.0 - eta)
    N3 = 0.25 * (1.0 + xi) * (1.0 + eta)
    N4 = 0.25 * (1.0 - xi) * (1.0 + eta)
    gx = (N1[None, :] * x0[:, None] + N2[None, :] * x1[:, None]
          + N3[None, :] * x1[:, None] + N4[None, :] * x0[:, None]).ravel()
    gy = (N1[None, :] * y0[:, None] + N2[None, :] * y0[:, None]
          + N3[None, :] * y1[:, None] + N4[None, :] * y1[:, None]).ravel()
    detJ = 0.25 * h * h
    w = np.tile(detJ * wx * wy, nc * nc)
    return GPSet(gx, gy, w)


def straight_crease_on_row(n_side: int, h: float, index: int, vertical: bool) -> Crease:
    """A crease through the node row/column `index`, boundary to boundary."""
    t = np.arange(n_side, dtype=np.float64) * h
    c = np.full(n_side, index * h)
    if vertical:
        return Crease(c.copy(), t.copy(), True, True, False)
    return Crease(t.copy(), c.copy(), True, True, False)


def arc_crease(cx: float, cy: float, r: float, a0: float, a1: float, nseg: int,
               ext_s: bool, ext_e: bool) -> Crease:
    a = a0 + (a1 - a0) * np.arange(nseg + 1, dtype=np.float64) / nseg
    return Crease(cx + r * np.cos(a), cy + r * np.sin(a), ext_s, ext_e, False)


def crease_gauss(cr: Crease) -> GPSet:
    """1-point rule per segment (Quadrature.h:99-121 with line_degree 1)."""
    p1x, p1y = cr.vx[:-1], cr.vy[:-1]
    p2x, p2y = cr.vx[1:], cr.vy[1:]
    tx = p2x - p1x; ty = p2y - p1y
    L = np.sqrt(tx * tx + ty * ty)
    keep = L > 0
    p1x, p1y, p2x, p2y, tx, ty, L = (a[keep] for a in (p1x, p1y, p2x, p2y, tx, ty, L))
    tx = tx / L; ty = ty / L
    nx = -ty; ny = tx
    gx = 0.5 * p1x + 0.5 * p2x
    gy = 0.5 * p1y + 0.5 * p2y
    w = (L / 2.0) * 2.0
    return GPSet(gx, gy, w, nx, ny, tx, ty)


def fold_from_crease(cr: Crease, Kt: float, Ang0: float, index: int) -> Fold:
    g = crease_gauss(cr)
    left = GPSet(g.gx + OFFSET * g.nx, g.gy + OFFSET * g.ny, g.w.copy(), g.nx, g.ny, g.tx, g.ty)
    right = GPSet(g.gx - OFFSET * g.nx, g.gy - OFFSET * g.ny, g.w.copy(), g.nx, g.ny, g.tx, g.ty)
    return Fold(left, right, Kt, Ang0, index)


def split_on_creases(face: GPSet, creases: List[Crease]):
    """plate/SetupHelper.h:60-106: a face GP within 1e-9 of a crease becomes two
    half-weight GPs at +/-1e-7 along the nearest segment's normal; the twins are
    appended in order."""
    ng = face.n
    best_d = np.full(ng, np.inf)
    best_nx = np.zeros(ng); best_ny = np.zeros(ng)
    # GPs sorted by x once: a segment's candidates (inside its box widened by
    # 1e-6) are a slice found by binary search, then filtered in y
    order = np.argsort(face.gx, kind="stable")
    sx = face.gx[order]
    for cr in creases:
        for i in range(len(cr.vx) - 1):
            ax, ay, bx, by = cr.vx[i], cr.vy[i], cr.vx[i + 1], cr.vy[i + 1]
            vx, vy = bx - ax, by - ay
            v2 = vx * vx + vy * vy
            if v2 == 0.0:
                continue
            lo_x, hi_x = min(ax, bx) - 1e-6, max(ax, bx) + 1e-6
            lo_y, hi_y = min(ay, by) - 1e-6, max(ay, by) + 1e-6
            cand = order[np.searchsorted(sx, lo_x, "left"):np.searchsorted(sx, hi_x, "right")]
            cand = np.sort(cand[(face.gy[cand] >= lo_y) & (face.gy[cand] <= hi_y)])
            if cand.size == 0:
                continue
            px, py = face.gx[cand], face.gy[cand]
            t = np.clip(((px - ax) * vx + (py - ay) * vy) / v2, 0.0, 1.0)
            d = np.hypot(px - (ax + t * vx), py - (ay + t * vy))
            better = d < best_d[cand]
            idx = cand[better]
            best_d[idx] = d[better]
            sv = math.sqrt(v2)
            best_nx[idx] = -vy / sv
            best_ny[idx] = vx / sv
    hit = np.nonzero(best_d < SPLIT_EPS)[0]
    if hit.size == 0:
        return face, 0
    gx, gy, w = face.gx.copy(), face.gy.copy(), face.w.copy()
    w[hit] *= 0.5
    tw_x = gx[hit] - OFFSET * best_nx[hit]
    tw_y = gy[hit] - OFFSET * best_ny[hit]
    gx[hit] += OFFSET * best_nx[hit]
    gy[hit] += OFFSET * best_ny[hit]
    out = GPSet(np.concatenate([gx, tw_x]), np.concatenate([gy, tw_y]),
                np.concatenate([w, w[hit]]))
    return out, int(hit.size)


def square_sheet(name: str, n_side: int, creases: List[Crease], Kts: List[float],
                 rho_factor: float = 2.5, order: int = 2, kernel: int = 0,
                 E: float = 1.0, nu: float = 0.3, t: float = 0.01, q: int = 4,
                 Ang0: float = 0.0) -> Sheet:
    X, Y, h = grid_nodes(n_side)
    face = face_gauss(n_side, h, q)
    face, nsplit = split_on_creases(face, creases)
    folds = [fold_from_crease(cr, kt, Ang0, i) for i, (cr, kt) in enumerate(zip(creases, Kts))]
    return Sheet(name, n_side, X, Y, face, creases, folds, rho_factor * h, order, kernel,
                 E, nu, t, nsplit)


# ---------------------------------------------------------------------------
# The configurations
# ---------------------------------------------------------------------------
def config1(kernel: int = 0, rho_factor: float = 2.5, n_side: int = 21) -> Sheet:
    """Unit square, uniform n x n grid, quadratic basis, 4x4 Gauss per cell,
    straight crease on the middle node column (x = 0.5), elastic hinge
    Kt = 1e-2 E t^3, E = 1, nu = 0.3, t = 0.01."""
    X, Y, h = grid_nodes(n_side)
    cr = straight_crease_on_row(n_side, h, (n_side - 1) // 2, vertical=True)
    E, t = 1.0, 0.01
    return square_sheet(f"straight_{n_side}x{n_side}", n_side, [cr], [1e-2 * E * t ** 3],
                        rho_factor=rho_factor, kernel=kernel)


ARC2_SEGMENTS = 256


def config2(kernel: int = 0, rho_factor: float = 2.5, n_side: int = 201,
            nseg: int = ARC2_SEGMENTS) -> Sheet:
    """201 x 201 nodes, curved crease: arc of radius 0.4 centred at (0.5, -0.1)
    from y = 0 to y = 0 (both ends on the boundary), as a polyline of `nseg`
    equal-angle segments."""
    a0 = math.asin(0.1 / 0.4)
    cr = arc_crease(0.5, -0.1, 0.4, a0, math.pi - a0, nseg, True, True)
    E, t = 1.0, 0.01
    return square_sheet(f"arc_{n_side}x{n_side}", n_side, [cr], [1e-2 * E * t ** 3],
                        rho_factor=rho_factor, kernel=kernel)


def config5(n_side: int = 1001, kernel: int = 0, rho_factor: float = 3.75, nseg_arc: int = 0) -> Sheet:
    """Large single sheet with a four-vertex-like crease layout: straight creases
    on the middle node column (x = 0.5, fold 0) and the middle node row (y = 0.5,
    fold 1), crossing at the centre, and a circular arc of radius 0.3 about the
    centre from 18 to 162 degrees (fold 2; interior tips, so the crease
    indicator's end caps are exercised). Fold 0 is meant to be the rigid hinge
    (constraint rows); the others are elastic, Kt = 1e-2 E t^3."""
    X, Y, h = grid_nodes(n_side)
    mid = (n_side - 1) // 2
    cv = straight_crease_on_row(n_side, h, mid, vertical=True)
    ch = straight_crease_on_row(n_side, h, mid, vertical=False)
    nseg = nseg_arc or max(32, (n_side - 1) // 2)
    arc = arc_crease(0.5, 0.5, 0.3, 0.1 * math.pi, 0.9 * math.pi, nseg, False, False)
    E, t = 1.0, 0.01
    Kt = 1e-2 * E * t ** 3
    return square_sheet(f"fourvertex_{n_side}x{n_side}", n_side, [cv, ch, arc], [Kt, Kt, Kt],
                        rho_factor=rho_factor, kernel=kernel)


BATCH_RHO_FACTOR = 3.75
BATCH_SHEETS = 256


def config_batched_sheet(s: int, seed: int = 3, kernel: int = 0) -> Sheet:
    """Sheet s of the batched set: side ~ U{21..121}, one straight crease on a
    node row U{3..n-4}, orientation 0 or 90 degrees, Kt ~ logU[1e-3, 1e-1] E t^3.
    Support 3.75 h: at 2.5 h the quadratic basis collapses next to a crease on a
    node row (the reference's pivot test fails there). Rows 1, 2, n-3, n-2 leave
    fewer than three visible node rows on the narrow side, where the quadratic
    moment matrix is singular (the reference's RKPM rejects such a sheet)."""
    u = uniform01(seed * 1000003 + s, 4)
    n = min(21 + int(u[0] * 101), 121)
    idx = 3 + min(int(u[1] * (n - 6)), n - 7)
    vertical = u[2] < 0.5
    E, t = 1.0, 0.01
    Kt = 10.0 ** (-3.0 + 2.0 * u[3]) * E * t ** 3
    X, Y, h = grid_nodes(n)
    cr = straight_crease_on_row(n, h, idx, vertical)
    return square_sheet(f"batch{s}_{n}", n, [cr], [Kt], rho_factor=BATCH_RHO_FACTOR, kernel=kernel)


def rigid_fold_displacement(sh: Sheet, angle: float, seed: int, sigma: float = 1e-3) -> np.ndarray:
    """Rigid fold of the + side of crease 0 (nodes with indicator f >= 0, which
    includes the nodes on the crease) by `angle` about the crease line, plus
    N(0, sigma) noise. Valid for straight creases."""
    cr = sh.creases[0]
    ax, ay = cr.vx[0], cr.vy[0]
    dx, dy = cr.vx[-1] - ax, cr.vy[-1] - ay
    L = math.hypot(dx, dy)
    tx, ty = dx / L, dy / L
    nx, ny = -ty, tx
    side = (sh.X - ax) * nx + (sh.Y - ay) * ny      # >= 0: + side
    plus = side >= -1e-12
    x = np.zeros(3 * sh.nn)
    d = side[plus]
    # rotate the + side about the crease axis by `angle`
    px = sh.X[plus] - d * nx
    py = sh.Y[plus] - d * ny
    newx = px + d * math.cos(angle) * nx
    newy = py + d * math.cos(angle) * ny
    newz = d * math.sin(angle)
    x[3 * np.nonzero(plus)[0] + 0] = newx - sh.X[plus]
    x[3 * np.nonzero(plus)[0] + 1] = newy - sh.Y[plus]
    x[3 * np.nonzero(plus)[0] + 2] = newz
    if sigma > 0:
        x += normal(seed, 3 * sh.nn, sigma)
    return x


def displacement(nn: int, seed: int = 1, sigma: float = 1e-3) -> np.ndarray:
    """u ~ N(0, sigma) per DOF (3 DOFs per node, node-major)."""
    return normal(seed, 3 * nn, sigma)


# ---------------------------------------------------------------------------
# Boundary domains and the Newton load case
# ---------------------------------------------------------------------------
def point_domain(x: float, y: float) -> GPSet:
    """0-D domain: one GP at the point, weight 1 (Quadrature.h:84-93)."""
    z = np.zeros(1)
    return GPSet(np.array([x]), np.array([y]), np.ones(1), z, z.copy(), z.copy(), z.copy())


def edge_domain(ax: float, ay: float, bx: float, by: float, nseg: int) -> GPSet:
    """Straight boundary edge split into `nseg` equal line elements, one GP per
    element at its midpoint, weight = element length, tangent along a->b,
    normal = tangent turned by +90 degrees (Quadrature.h:96-121, line_degree 1)."""
    s0 = np.arange(nseg) / nseg
    s1 = (np.arange(nseg) + 1) / nseg
    p1x, p1y = ax + (bx - ax) * s0, ay + (by - ay) * s0
    p2x, p2y = ax + (bx - ax) * s1, ay + (by - ay) * s1
    tx, ty = p2x - p1x, p2y - p1y
    L = np.sqrt(tx * tx + ty * ty)
    tx, ty = tx / L, ty / L
    gx = 0.5 * p1x + 0.5 * p2x
    gy = 0.5 * p1y + 0.5 * p2y
    return GPSet(gx, gy, L.copy(), -ty, tx.copy(), tx, ty)


def bending_d11(sh: Sheet) -> float:
    return sh.E * sh.t ** 3 / (12.0 * (1.0 - sh.nu ** 2))


def with_edge_moment_load(sh: Sheet, seed: int = 2) -> Sheet:
    """The Newton load case: the three-corner supports of the reference's
    regression configurations (point 1 at (0,0): u, v, w; point 2 at (1,0): v, w;
    point 4 at (0,1): w) and equal normal edge moments on x = 0 and x = 1.
    The moments are self-equilibrated, so the supports only remove the rigid
    motions. Magnitude m = s Kt, s = 0.5 + U(0,1) from SplitMix64(seed), Kt the
    hinge stiffness per unit length of crease 0: a straight crease then opens by
    about s radians at full load, while the facets bend to a curvature m / D of
    about 0.1 s (D = E t^3 / 12 (1 - nu^2)). Without a crease, m = s D."""
    n = sh.n_side
    s = 0.5 + float(uniform01(seed, 1)[0])
    m = s * (sh.folds[0].Kt if sh.folds else bending_d11(sh))
    b = [Boundary(point_domain(0.0, 0.0), "point", active=1 | 2 | 4),
         Boundary(point_domain(1.0, 0.0), "point", active=2 | 4),
         Boundary(point_domain(0.0, 1.0), "point", active=4),
         Boundary(edge_domain(0.0, 0.0, 0.0, 1.0, n - 1), "edge", moment=m),
         Boundary(edge_domain(1.0, 0.0, 1.0, 1.0, n - 1), "edge", moment=m)]
    return dataclasses.replace(sh, name=sh.name + "_moment", boundaries=b)


def newton_tolerances(f_ext_inf: float) -> dict:
    """Solver tolerances for the load path, scaled to the load: the reference's
    defaults (tol_force 1e-8, tol_energy 1e-8) are absolute and its examples set
    them per configuration (solver.tol_force / solver.tol_energy, main.cpp:263-264);
    with E = 1, t = 0.01 the loads are ~1e-8, so they are set relative to
    max|F_ext|."""
    return {"tol_force": 1e-7 * f_ext_inf, "tol_energy": 1e-9 * f_ext_inf}


NEWTON_INCREMENTS = 20


# ---------------------------------------------------------------------------
# Serving requests
# ---------------------------------------------------------------------------
SERVING_RHO_FACTOR = 3.75


def serving_sheet(n_side: int, kernel: int = 0) -> Sheet:
    """Sheet of a serving request: n x n nodes on the unit square, straight
    crease on the node column nearest the centre ((n - 1) // 2), elastic hinge
    1e-2 E t^3, quadratic RKPM with support 3.75 h."""
    X, Y, h = grid_nodes(n_side)
    cr = straight_crease_on_row(n_side, h, (n_side - 1) // 2, vertical=True)
    E, t = 1.0, 0.01
    return square_sheet(f"serve_{n_side}", n_side, [cr], [1e-2 * E * t ** 3],
                        rho_factor=SERVING_RHO_FACTOR, kernel=kernel)


def serving_requests(nreq: int, seed: int = 4, n_min: int = 21, n_max: int = 201):
    """(side, fold angle in radians, noise seed) per request: side ~ U{n_min..n_max},
    angle ~ U[0, 170 deg]."""
    u = uniform01(seed, 2 * nreq).reshape(nreq, 2)
    sides = np.minimum(n_min + (u[:, 0] * (n_max - n_min + 1)).astype(int), n_max)
    angles = u[:, 1] * math.radians(170.0)
    return [(int(sides[i]), float(angles[i]), seed * 1000003 + i) for i in range(nreq)]


def poisson_arrivals(rate: float, n: int, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson stream: exponential gaps -ln(1 - U) / rate."""
    u = uniform01(seed, n)
    return np.cumsum(-np.log1p(-u) / rate)
