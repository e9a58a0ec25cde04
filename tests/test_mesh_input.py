"""Mesh input (Gmsh reader, quadrature, crease polylines, support radius):
the library's host entry points against the oracle's restatement, and the
oracle pinned by what the meshes' geometry and the Gauss rules fix.

Bars: bit-exact (node coordinates, Gauss points, weights, frames, crease
vertices, rho), equal flags and counts. No GPU is needed: these entry points
do no device work.

Fixtures (tests/golden/reference_inputs/): the reference's own meshes and
configurations (examples/smoke_centre_crease/CB_n20.msh and smoke.json,
tests/regression/meshes/*.msh, tests/regression/configs/*.json).
"""
import math
import os

import numpy as np
import pytest

import curvami_gpu as G
import mesh_setup as MS
import mesh_oracle as OM

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "reference_inputs")
MESHES = ["CB_n20", "Arc_LR0p4_Phi30", "Miura", "SquareTwist"]


def mpath(name):
    return os.path.join(INPUTS, name + ".msh")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


@pytest.fixture(scope="module", params=MESHES)
def pair(request):
    return request.param, G.Mesh(mpath(request.param)), OM.read_gmsh(mpath(request.param))


def _header_counts(path):
    lines = open(path).read().split("\n")
    n = lines[lines.index("$Nodes") + 1].split()
    e = lines[lines.index("$Elements") + 1].split()
    return int(n[1]), int(e[1])


def test_reader_matches_oracle_and_file(pair):
    name, m, o = pair
    nn, ne = _header_counts(mpath(name))
    assert m.X.size == nn == len(o.X) and m.type.size == ne == len(o.elements)
    assert np.array_equal(bits(m.X), bits(o.X)) and np.array_equal(bits(m.Y), bits(o.Y))
    assert m.type.tolist() == [e[0] for e in o.elements]
    assert m.entity.tolist() == [e[1] for e in o.elements]
    idx = [m.node_idx[m.node_off[k]:m.node_off[k + 1]].tolist() for k in range(m.type.size)]
    assert idx == [e[2] for e in o.elements]


def test_cb_n20_matches_its_geometry():
    """CB_n20: [-1, 1]^2 with 21 x 21 transfinite nodes, 20 x 20 quads, a crease
    line (-0.5, 0) -> (0.5, 0) of 10 elements, corner points tagged 1, 2, 4."""
    m = G.Mesh(mpath("CB_n20"))
    q = m.quadrature(1, 3)
    X, Y = m.X[q.active], m.Y[q.active]
    assert q.active.size == 441 and q.face.shape[0] == 400 * 9
    g = np.linspace(-1.0, 1.0, 21)
    assert np.abs(np.sort(np.unique(np.round(X, 9))) - g).max() < 1e-9
    assert np.abs(np.sort(np.unique(np.round(Y, 9))) - g).max() < 1e-9
    assert sorted(q.point_tag.tolist()) == [1, 2, 4] and set(q.edge_tag.tolist()) == {5}
    assert abs(q.edge[:, 2].sum() - 1.0) < 1e-12
    vx, vy, closed, es, ee, dropped = m.extract_curve([5])
    assert vx.size == 11 and not closed and not es and not ee and dropped == 0
    assert abs(vx[0] + 0.5) < 1e-12 and abs(vx[-1] - 0.5) < 1e-12 and np.abs(vy).max() < 1e-12
    rho = G.support_radius(X, Y, 3)
    assert abs(rho - 0.6) < 1e-11          # (order + 3) h with h = 0.1


@pytest.mark.parametrize("degrees", [(1, 3), (3, 2), (1, 1)])
def test_quadrature_matches_oracle(pair, degrees):
    name, m, o = pair
    q = m.quadrature(*degrees)
    r = OM.quadrature(o, *degrees)
    assert np.array_equal(q.active, r["active"])
    for k in ("face", "edge", "edge_tn", "point"):
        a = getattr(q, k)
        assert a.shape == r[k].shape, k
        assert np.array_equal(bits(a), bits(r[k])), f"{name} {degrees}: {k} not bit-identical"
    for k in ("face_tag", "edge_tag", "point_tag"):
        assert np.array_equal(getattr(q, k), r[k]), k


def _box_integral(a, b, x0, x1, y0, y1):
    return (x1 ** (a + 1) - x0 ** (a + 1)) / (a + 1) * (y1 ** (b + 1) - y0 ** (b + 1)) / (b + 1)


@pytest.mark.parametrize("name", ["CB_n20", "Miura", "SquareTwist"])
def test_quad_rule_integrates_degree5_exactly(name):
    """On the square sheets the 3 x 3 rule integrates x^a y^b (a, b <= 5) exactly
    (structured meshes: affine quads), so the oracle's points and weights are
    pinned by closed-form integrals, not by its own formulas."""
    o = OM.read_gmsh(mpath(name))
    r = OM.quadrature(o, 1, 3)
    f = r["face"]
    X = np.array(o.X)[r["active"]]; Y = np.array(o.Y)[r["active"]]
    x0, x1, y0, y1 = X.min(), X.max(), Y.min(), Y.max()
    for a in range(6):
        for b in range(6):
            exact = _box_integral(a, b, x0, x1, y0, y1)
            got = float(np.sum(f[:, 2] * f[:, 0] ** a * f[:, 1] ** b))
            assert abs(got - exact) <= 1e-12 * max(1.0, abs(exact)) * 50, (a, b, got, exact)
    # the 2 x 2 rule: exact to degree 3 per variable, not 4 in general
    r2 = OM.quadrature(o, 1, 2)["face"]
    assert abs(np.sum(r2[:, 2] * r2[:, 0] ** 3 * r2[:, 1] ** 2) - _box_integral(3, 2, x0, x1, y0, y1)) < 1e-12


def _write_msh(path, nodes, blocks):
    """nodes: list of (tag, x, y); blocks: list of (dim, entity, type, [(tag, [node tags])])."""
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n")
        f.write(f"1 {len(nodes)} 1 {max(t for t, _, _ in nodes)}\n2 1 0 {len(nodes)}\n")
        for t, _, _ in nodes:
            f.write(f"{t}\n")
        for _, x, y in nodes:
            f.write(f"{x!r} {y!r} 0\n")
        ne = sum(len(b[3]) for b in blocks)
        f.write(f"$EndNodes\n$Elements\n{len(blocks)} {ne} 1 {ne}\n")
        for dim, ent, ty, els in blocks:
            f.write(f"{dim} {ent} {ty} {len(els)}\n")
            for et, nv in els:
                f.write(f"{et} " + " ".join(map(str, nv)) + "\n")
        f.write("$EndElements\n")


def test_triangle_and_line_rules(tmp_path):
    """Triangles: the 3-point rule integrates quadratics exactly, the centroid
    rule linears; 3-point line rule: degree 5 along a segment."""
    P = [(1, 0.0, 0.0), (2, 1.3, 0.2), (3, 0.4, 1.1), (4, 1.5, 1.4)]
    tris = [(1, [1, 2, 3]), (2, [2, 4, 3])]
    lines = [(3, [1, 4])]
    p = str(tmp_path / "tri.msh")
    _write_msh(p, P, [(2, 7, 2, tris), (1, 9, 1, lines)])
    m = G.Mesh(p)
    o = OM.read_gmsh(p)
    XY = {t: (x, y) for t, x, y in P}

    def tri_int(a, b, c, f):
        """Edge-midpoint rule: exact for quadratics, and a different rule from
        the interior 3-point one under test."""
        A = 0.5 * abs((b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]))
        mids = [((a[0] + b[0]) / 2, (a[1] + b[1]) / 2), ((b[0] + c[0]) / 2, (b[1] + c[1]) / 2),
                ((c[0] + a[0]) / 2, (c[1] + a[1]) / 2)]
        return A / 3 * sum(f(*q) for q in mids)

    for deg, polys in ((3, [(2, 0), (1, 1), (0, 2), (1, 0)]), (1, [(1, 0), (0, 1), (0, 0)])):
        q = m.quadrature(3, deg)
        r = OM.quadrature(o, 3, deg)
        assert np.array_equal(bits(q.face), bits(r["face"])) and np.array_equal(bits(q.edge), bits(r["edge"]))
        for a, b in polys:
            f = lambda x, y: x ** a * y ** b
            exact = sum(tri_int(XY[t[0]], XY[t[1]], XY[t[2]], f) for _, t in tris)
            got = float(np.sum(q.face[:, 2] * q.face[:, 0] ** a * q.face[:, 1] ** b))
            assert abs(got - exact) < 1e-13, (deg, a, b)
    q = m.quadrature(3, 3)
    (x0, y0), (x1, y1) = XY[1], XY[4]
    L = math.hypot(x1 - x0, y1 - y0)
    for k in range(6):     # integral of s^k along the segment, s = arc length
        s = np.hypot(q.edge[:, 0] - x0, q.edge[:, 1] - y0)
        assert abs(np.sum(q.edge[:, 2] * s ** k) - L ** (k + 1) / (k + 1)) < 1e-12
    assert np.allclose(q.edge_tn[0], [(x1 - x0) / L, (y1 - y0) / L, -(y1 - y0) / L, (x1 - x0) / L])
    assert q.active.tolist() == [0, 1, 2, 3] and q.point.shape[0] == 0


def test_convex_hull_oracle_against_scipy():
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(11)
    for n in (5, 40, 500):
        pts = rng.random((n, 2))
        h = OM.convex_hull([tuple(p) for p in pts])
        ref = pts[ConvexHull(pts).vertices]
        assert sorted(map(tuple, h)) == sorted(map(tuple, ref))


def test_extract_curve_matches_oracle():
    for cfg in ("smoke", "ls_miura", "tdal_squaretwist", "al_arccrease"):
        import json
        c = json.load(open(os.path.join(INPUTS, cfg + ".json")))
        name = os.path.splitext(os.path.basename(c["mesh_file"]))[0]
        m = G.Mesh(mpath(name)); o = OM.read_gmsh(mpath(name))
        groups = c.get("cuts") or [[f["crease_tag"]] for f in c["folds"]]
        for tags in groups:
            got = m.extract_curve(tags)
            ref = OM.extract_curve(o, tags)
            assert np.array_equal(bits(got[0]), bits(ref[0])) and np.array_equal(bits(got[1]), bits(ref[1]))
            assert got[2:] == ref[2:], (cfg, tags, got[2:], ref[2:])


def test_extract_curve_boundary_flags_and_loops(tmp_path):
    """A crease from boundary to boundary extends at both ends; one ending
    inside does not; a loop is closed and repeats its first vertex; a branch
    drops nodes (and the sheet builder refuses it)."""
    nodes = [(k + 1, float(x), float(y)) for k, (x, y) in enumerate(
        [(0, 0), (4, 0), (4, 4), (0, 4), (1, 2), (2, 2), (3, 2), (4, 2), (2, 1), (2, 3), (1, 1), (3, 1), (3, 3),
         (1, 3)])]
    quad = [(1, [1, 2, 3, 4])]
    straight = [(2, [5, 6]), (3, [6, 7]), (4, [7, 8])]          # interior -> right edge
    loop = [(5, [11, 12]), (6, [12, 13]), (7, [13, 14]), (8, [14, 11])]
    branch = [(9, [6, 9]), (10, [6, 10])]                          # with the straight line: a T
    p = str(tmp_path / "c.msh")
    _write_msh(p, nodes, [(2, 1, 3, quad), (1, 5, 1, straight), (1, 6, 1, loop), (1, 7, 1, branch)])
    m = G.Mesh(p)
    o = OM.read_gmsh(p)
    vx, vy, closed, es, ee, dr = m.extract_curve([5])
    assert vx.tolist() == [1.0, 2.0, 3.0, 4.0] and (closed, es, ee, dr) == (False, False, True, 0)
    vx, vy, closed, es, ee, dr = m.extract_curve([6])
    assert closed and vx.size == 5 and (vx[0], vy[0]) == (vx[-1], vy[-1]) and dr == 0
    got = m.extract_curve([5, 7])
    ref = OM.extract_curve(o, [5, 7])
    assert got[5] > 0 and got[5] == ref[5] and np.array_equal(got[0], ref[0])
    with pytest.raises(ValueError):
        MS.crease_from_mesh(m, [5, 7])
    with pytest.raises(ValueError):
        MS.crease_from_mesh(m, [42])


def test_support_radius_matches_oracle(pair):
    name, m, o = pair
    q = m.quadrature(1, 3)
    X, Y = m.X[q.active], m.Y[q.active]
    for order in (2, 3):
        a = G.support_radius(X, Y, order)
        b = OM.support_radius(X, Y, order)
        assert np.float64(a).view(np.int64) == np.float64(b).view(np.int64), (name, order, a, b)


def test_support_radius_closed_form_and_scattered():
    h = 0.125
    g = np.arange(33) * h
    X, Y = np.meshgrid(g, g)
    X, Y = X.ravel(), Y.ravel()
    assert G.support_radius(X, Y, 3) == 6 * h and G.support_radius(X, Y, 2) == 5 * h
    assert G.support_radius(X[:3], Y[:3], 3) == 0.1
    rng = np.random.default_rng(3)
    for n in (4, 5, 100, 3000):
        P = rng.random((n, 2)) * [3.0, 0.2]   # elongated cloud: many empty cells
        a = G.support_radius(P[:, 0], P[:, 1], 3)
        assert a == OM.support_radius(P[:, 0], P[:, 1], 3)


def test_reader_errors(tmp_path):
    good = [(1, 0.0, 0.0), (2, 1.0, 0.0), (3, 1.0, 1.0), (4, 0.0, 1.0)]
    p = str(tmp_path / "bad_type.msh")
    _write_msh(p, good, [(2, 1, 9, [(1, [1, 2, 3, 4, 1, 2])])])       # 6-node triangle
    with pytest.raises(G.CvmError, match="unsupported Gmsh element type 9"):
        G.Mesh(p)
    p = str(tmp_path / "bad_tag.msh")
    _write_msh(p, good, [(2, 1, 3, [(1, [1, 2, 3, 99])])])
    with pytest.raises(G.CvmError, match="undefined node tag 99"):
        G.Mesh(p)
    p = str(tmp_path / "no_elems.msh")
    _write_msh(p, good, [(2, 1, 3, [])])
    with pytest.raises(G.CvmError, match="empty or malformed"):
        G.Mesh(p)
    with pytest.raises(G.CvmError, match="could not open"):
        G.Mesh(str(tmp_path / "missing.msh"))
    p = str(tmp_path / "binary.msh")
    with open(p, "w") as f:
        f.write("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    with pytest.raises(G.CvmError, match="binary"):
        G.Mesh(p)


def test_sheet_from_config_mirrors_driver_setup():
    """smoke.json on CB_n20: cubic basis, rho = 6h, one elastic fold on tag 5
    (GPs at +/-1e-7 along the crease normal), the crease from the fold's tag,
    point domains 1, 2, 4 with u v w / v w / w fixed, the crease edge domain."""
    ms = MS.sheet_from_config(os.path.join(INPUTS, "smoke.json"), mesh_file=mpath("CB_n20"))
    sh = ms.sheet
    assert sh.order == 3 and abs(sh.rho - 0.6) < 1e-11 and sh.nn == 441 and sh.face.n == 3600
    assert (sh.E, sh.nu, sh.t) == (96000000.0, 0.3, 0.005)
    assert len(sh.folds) == 1 and len(sh.creases) == 1 and not sh.folds[0].rigid
    f = sh.folds[0]
    e = ms.input.edges[5]
    assert np.array_equal(f.left.gy, e.gy + MS.OFFSET * e.ny) and np.array_equal(f.right.gy, e.gy - MS.OFFSET * e.ny)
    assert abs(f.Kt - 1.098901098901099) < 1e-15 and f.Ang0 == 0.0
    kinds = [(b.kind, b.active) for b in sh.boundaries]
    assert kinds == [("point", 7), ("point", 6), ("point", 4), ("edge", 0)]
    ms2 = MS.sheet_from_config(os.path.join(INPUTS, "ls_miura.json"), mesh_file=mpath("Miura"))
    assert [fl.rigid for fl in ms2.sheet.folds] == [True] * 4 and len(ms2.sheet.creases) == 2
    assert abs(ms2.sheet.folds[0].Ang0 - (-15.26158022631557 * math.pi / 180.0)) < 1e-15


def test_natural_curvature_config():
    """The "natural_curvature" band of a configuration (main.cpp:80-97): degrees to
    radians, mountain -1 / valley +1 (mountain by default), the reference's errors."""
    import mesh_setup as MS
    cfg = {"natural_curvature": {"type": "band", "x0": 0.25, "width": 0.1, "angle_deg": 45.0}}
    assert MS.natural_curvature_from_config(cfg) == (0.25, 0.1, math.pi / 4, -1.0)
    cfg["natural_curvature"]["sense"] = "valley"
    assert MS.natural_curvature_from_config(cfg)[3] == 1.0
    assert MS.natural_curvature_from_config({}) is None
    for bad, msg in (({"type": "disk"}, "natural_curvature.type must be band"),
                     ({"sense": "ridge"}, "natural_curvature.sense must be mountain or valley"),
                     ({"width": 0.0}, "natural_curvature.width must be > 0")):
        c = {"natural_curvature": {"type": "band", "x0": 0.0, "width": 0.1, "angle_deg": 10.0, **bad}}
        with pytest.raises(ValueError, match=msg):
            MS.natural_curvature_from_config(c)
