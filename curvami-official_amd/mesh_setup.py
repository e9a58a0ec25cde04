"""Creased sheets from the reference's own inputs: a Gmsh mesh plus a JSON
configuration, set up as the reference's driver does before its solve.

Follows
  SetupHelper::loadMeshAndQuadrature   plate/SetupHelper.h:14-52
  SetupHelper::splitGaussPointsOnCreases plate/SetupHelper.h:60-106
  main.cpp:50-78   material (one isotropic ply), mesh, rho = computeSupportRadius(nodes, 3)
  main.cpp:104-176 folds (Meshless::addFold, plate/Meshless.h:30-46) and the
                   crease polylines from "cuts" or from each fold's crease tag
  main.cpp:209-252 point and edge boundary conditions
The mesh reading, quadrature, crease extraction and support radius run in the
library (cvm_mesh_*, cvm_support_radius); this module arranges their arrays
into a workloads.Sheet, the input of the device plate.

The reference's driver builds a cubic basis (Meshless plate(3), main.cpp:53);
`order` defaults to 3 here. Analysis settings (solver, creasing, outputs) are
returned untouched in MeshSheet.config for the caller.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from typing import Dict, List, Optional

import numpy as np

import curvami_gpu as G
from workloads import OFFSET, Boundary, Crease, Fold, GPSet, Sheet, split_on_creases

LINE_DEGREE = 1   # plate/SetupHelper.h:28-30: Quadrature::generate(..., 1, 3)
FACE_DEGREE = 3


@dataclasses.dataclass
class MeshInput:
    """Mesh + quadrature, grouped like the reference's integration domains:
    face GPs, point and edge domains keyed by entity tag in ascending order
    (std::map, plate/SetupHelper.h:32-50)."""
    mesh: "G.Mesh"
    X: np.ndarray                  # active nodes (face-element nodes, ascending raw index)
    Y: np.ndarray
    active: np.ndarray             # raw index of each active node
    face: GPSet
    face_tag: np.ndarray
    points: Dict[int, GPSet]
    edges: Dict[int, GPSet]


def load_mesh_and_quadrature(path: str, line_degree: int = LINE_DEGREE,
                             face_degree: int = FACE_DEGREE) -> MeshInput:
    mesh = G.Mesh(path)
    q = mesh.quadrature(line_degree, face_degree)
    face = GPSet(q.face[:, 0].copy(), q.face[:, 1].copy(), q.face[:, 2].copy())
    points: Dict[int, GPSet] = {}
    for tag in sorted(set(q.point_tag.tolist())):
        k = np.nonzero(q.point_tag == tag)[0]
        points[tag] = GPSet(q.point[k, 0].copy(), q.point[k, 1].copy(), q.point[k, 2].copy())
    edges: Dict[int, GPSet] = {}
    for tag in sorted(set(q.edge_tag.tolist())):
        k = np.nonzero(q.edge_tag == tag)[0]
        tn = q.edge_tn[k]
        edges[tag] = GPSet(q.edge[k, 0].copy(), q.edge[k, 1].copy(), q.edge[k, 2].copy(),
                           nx=tn[:, 2].copy(), ny=tn[:, 3].copy(), tx=tn[:, 0].copy(), ty=tn[:, 1].copy())
    act = q.active.astype(np.int64)
    return MeshInput(mesh, mesh.X[act].copy(), mesh.Y[act].copy(), act, face, q.face_tag, points, edges)


def crease_from_mesh(mesh: "G.Mesh", tags) -> Crease:
    """One crease polyline (Enrichment::FoldLine) from line elements
    (MeshUtils::extractCurve). Raises if the tags give no line elements."""
    vx, vy, closed, ext_s, ext_e, dropped = mesh.extract_curve(tags)
    if vx.size == 0:
        raise ValueError(f"no line elements with entity tags {list(np.atleast_1d(tags))}")
    if dropped:
        raise ValueError(f"crease tags {list(np.atleast_1d(tags))}: {dropped} nodes are not on one chain "
                         "(disconnected or branching crease groups are not supported; split them into cuts)")
    return Crease(vx, vy, ext_s, ext_e, closed)


def fold_from_edge(edge: GPSet, Kt: float, Ang0: float, index: int, rigid: bool = False,
                   force_left=(0.0, 0.0, 0.0), force_right=(0.0, 0.0, 0.0)) -> Fold:
    """Meshless::addFold (plate/Meshless.h:30-46): the edge domain's GPs moved
    by +/-1e-7 along their normal, weights and frames kept."""
    left = GPSet(edge.gx + OFFSET * edge.nx, edge.gy + OFFSET * edge.ny, edge.w.copy(),
                 edge.nx.copy(), edge.ny.copy(), edge.tx.copy(), edge.ty.copy())
    right = GPSet(edge.gx - OFFSET * edge.nx, edge.gy - OFFSET * edge.ny, edge.w.copy(),
                  edge.nx.copy(), edge.ny.copy(), edge.tx.copy(), edge.ty.copy())
    return Fold(left, right, Kt, Ang0, index, rigid, tuple(force_left), tuple(force_right))


@dataclasses.dataclass
class FoldInfo:
    crease_tag: int
    angle_start: float    # rad, unsigned
    angle_target: float   # rad, unsigned
    sign: float           # +1 valley, -1 mountain
    is_driver: bool


@dataclasses.dataclass
class MeshSheet:
    sheet: Sheet
    input: MeshInput
    fold_info: List[FoldInfo]
    config: dict


def _apply_bc(bd: Boundary, cfg: dict, edge: bool):
    """BoundaryCondition fields from one point_bcs / edge_bcs entry (main.cpp:209-249)."""
    active = bd.active
    val = list(bd.val)
    for bit, key in enumerate(("u", "v", "w") + (("rot",) if edge else ())):
        if key in cfg:
            active |= 1 << bit
            val[bit] = float(cfg[key])
    bd.active = active
    bd.val = tuple(val)
    if "force" in cfg:
        bd.force = tuple(float(v) for v in cfg["force"])
    if edge and "moment" in cfg:
        bd.moment = float(cfg["moment"])


def natural_curvature_from_config(config: dict):
    """The configuration's "natural_curvature" band (main.cpp:80-97) as
    (x0, width, theta0 [rad], sign), sign -1 for a mountain and +1 for a valley;
    None without one. Raises ValueError with the reference's messages."""
    if "natural_curvature" not in config:
        return None
    nc = config["natural_curvature"]
    if nc.get("type", "band") != "band":
        raise ValueError("natural_curvature.type must be band")
    x0 = float(nc["x0"])
    width = float(nc["width"])
    theta0 = float(nc["angle_deg"]) * math.pi / 180.0
    sense = nc.get("sense", "mountain")
    if sense not in ("mountain", "valley"):
        raise ValueError("natural_curvature.sense must be mountain or valley")
    sign = -1.0 if sense == "mountain" else 1.0
    if not (width > 0.0):
        raise ValueError("natural_curvature.width must be > 0")
    return (x0, width, theta0, sign)


def sheet_from_config(config, base_dir: Optional[str] = None, order: int = 3, kernel: int = 0,
                      mesh_file: Optional[str] = None) -> MeshSheet:
    """Sheet for a reference configuration (a dict or a JSON path). mesh_file
    (taken as given) overrides config["mesh_file"], which resolves against
    base_dir (default: the JSON file's directory, else the working directory)."""
    if isinstance(config, str):
        if base_dir is None:
            base_dir = os.path.dirname(os.path.abspath(config))
        with open(config) as f:
            config = json.load(f)
    base_dir = base_dir or os.getcwd()
    path = mesh_file or config["mesh_file"]
    if mesh_file is None and not os.path.isabs(path):
        path = os.path.join(base_dir, path)
    natural = natural_curvature_from_config(config)
    mat = config["material"]
    inp = load_mesh_and_quadrature(path)
    rho = G.support_radius(inp.X, inp.Y, 3)            # main.cpp:78

    folds: List[Fold] = []
    info: List[FoldInfo] = []
    for fc in config.get("folds", []):                 # main.cpp:104-168
        tag = int(fc["crease_tag"])
        if "tie" in fc and fc["tie"].get("type", "pointwise") != "pointwise":
            raise NotImplementedError("only pointwise fold ties are supported")
        sign = -1.0 if fc.get("type", "") in ("mountain", "Mountain") else 1.0
        a0 = float(fc.get("angle_start", 0.0)) * math.pi / 180.0
        a1 = float(fc.get("angle_target", 0.0)) * math.pi / 180.0
        if tag not in inp.edges:
            raise ValueError(f"crease tag {tag} not found among the mesh's edge domains")
        folds.append(fold_from_edge(inp.edges[tag], float(fc.get("stiffness", 0.0)), a0 * sign, len(folds),
                                    bool(fc["is_rigid"]), fc.get("force_left", (0.0, 0.0, 0.0)),
                                    fc.get("force_right", (0.0, 0.0, 0.0))))
        info.append(FoldInfo(tag, a0, a1, sign, bool(fc.get("is_driver", False))))

    if "cuts" in config:                               # main.cpp:170-184
        creases = [crease_from_mesh(inp.mesh, tags) for tags in config["cuts"]]
    else:
        creases = [crease_from_mesh(inp.mesh, [int(fc["crease_tag"])]) for fc in config.get("folds", [])]

    face, n_split = split_on_creases(inp.face, creases)   # SetupHelper.h:140-143

    boundaries: List[Boundary] = []
    for tag, g in inp.points.items():                  # main.cpp:209-226
        bd = Boundary(g, "point")
        for bc in config.get("point_bcs", []):
            if int(bc["tag"]) == tag:
                _apply_bc(bd, bc, edge=False)
        boundaries.append(bd)
    for tag, g in inp.edges.items():                   # main.cpp:228-249
        bd = Boundary(g, "edge")
        for bc in config.get("edge_bcs", []):
            if int(bc["tag"]) == tag:
                _apply_bc(bd, bc, edge=True)
        boundaries.append(bd)

    name = os.path.splitext(os.path.basename(path))[0]
    sheet = Sheet(name, 0, inp.X, inp.Y, face, creases, folds, rho, order, kernel,
                  E=float(mat["E"]), nu=float(mat["nu"]), t=float(mat["h"]), n_split=n_split,
                  boundaries=boundaries, natural_curvature=natural)
    return MeshSheet(sheet, inp, info, config)
