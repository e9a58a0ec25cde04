"""ctypes binding of the CPU oracle (oracle.cpp). TEST INFRASTRUCTURE ONLY:
imported by tests/, by __graft_entry__.smoke() as the checker, and by the CPU
baseline leg of bench.py. The product library never uses it."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_c = C.POINTER(C.c_char)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or \
            os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "oracle.cpp")):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        sig = {
            "orc_create": (P, []),
            "orc_destroy": (None, [P]),
            "orc_last_error": (C.c_char_p, [P]),
            "orc_set_nodes": (None, [P, C.c_int, _d, _d]),
            "orc_set_material_plies": (None, [P, C.c_int, _d]),
            "orc_get_abd": (None, [P, _d, _d, _d]),
            "orc_add_crease": (None, [P, C.c_int, _d, _d, C.c_int, C.c_int, C.c_int]),
            "orc_add_domain": (C.c_int, [P, C.c_int, _d, _d, _d, _d, _d, _d, _d, C.c_int]),
            "orc_add_fold": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_double]),
            "orc_set_fold_gp_state": (None, [P, C.c_int, C.c_int, _d, _c]),
            "orc_set_fold_rigid": (None, [P, C.c_int, C.c_int]),
            "orc_indicator": (None, [P, C.c_int, C.c_double, C.c_double, _d]),
            "orc_window": (None, [C.c_int, C.c_int, _d, _d, _d, _d]),
            "orc_neighbours": (C.c_long, [P, C.c_int, C.c_double]),
            "orc_get_neighbours": (None, [P, C.c_int, _i, _i, _d]),
            "orc_set_neighbours": (None, [P, C.c_int, C.c_int, _i, _i, _d]),
            "orc_shape_functions": (C.c_int, [P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _i, _i]),
            "orc_get_shape_functions": (None, [P, C.c_int, _d, _d, _d, _d, _d, _d]),
            "orc_init_assembly": (C.c_int, [P]),
            "orc_face_nnz": (C.c_long, [P]),
            "orc_get_face_pattern": (None, [P, _i, _i]),
            "orc_get_gp_map": (None, [P, C.c_int, _i]),
            "orc_assemble": (C.c_long, [P, _d, C.c_int, C.c_int, C.c_int]),
            "orc_internal_force": (None, [P, _d, C.c_int, C.c_int, C.c_int, _d]),
            "orc_get_system": (None, [P, _i, _i, _d, _d]),
            "orc_last_seconds": (C.c_double, [P]),
            "orc_fold_angle": (C.c_double, [_d, C.c_double, C.c_double, _d]),
            "orc_kinematics": (None, [P, C.c_int, _d, _d]),
            "orc_set_domain_bc": (None, [P, C.c_int, C.c_int, C.c_int, _d, _d, C.c_double]),
            "orc_constraints": (C.c_long, [P, _d, C.c_int, C.POINTER(C.c_long)]),
            "orc_get_constraints": (None, [P, _d, _i, _i, _d]),
            "orc_external_force": (None, [P, _d]),
            "orc_energies": (None, [P, _d, _d]),
            "orc_set_natural_curvature": (None, [P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_fold_angles": (None, [P, C.c_int, _d, _d]),
            "orc_clear_creases": (None, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_i) if a is not None else None


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def window(kernel: int, s) -> tuple:
    s = _f64(s)
    w = np.empty_like(s); dw = np.empty_like(s); d2w = np.empty_like(s)
    lib().orc_window(kernel, s.size, _dp(s), _dp(w), _dp(dw), _dp(d2w))
    return w, dw, d2w


def fold_angle(F, tx, ty):
    F = _f64(F)
    g = np.empty(12)
    th = lib().orc_fold_angle(_dp(F), tx, ty, _dp(g))
    return th, g


class OraclePlate:
    """Mirror of Meshless (plate/Meshless.h) on the CPU oracle."""

    FACE = 0

    def __init__(self):
        self._L = lib()
        self._p = self._L.orc_create()
        self.domains = []
        self.ndoms = 0
        self.nn = 0

    def __del__(self):
        try:
            if self._p:
                self._L.orc_destroy(self._p)
        except Exception:
            pass

    # --- setup -----------------------------------------------------------
    def set_nodes(self, X, Y):
        X, Y = _f64(X), _f64(Y)
        self.nn = X.size
        self._keep_nodes = (X, Y)
        self._L.orc_set_nodes(self._p, X.size, _dp(X), _dp(Y))

    def set_plies(self, plies):
        plies = _f64(plies)
        self._L.orc_set_material_plies(self._p, plies.shape[0], _dp(plies))

    def set_natural_curvature(self, active=True, x0=0.0, width=1.0, theta0=0.0, sign=-1.0):
        """Kinematics::naturalCurvature() (plate/Kinematics.h:20-45): the band
        sign * theta0 * g(X) in E1_el(0, 0); sign -1 is a mountain, +1 a valley."""
        self._L.orc_set_natural_curvature(self._p, int(active), float(x0), float(width), float(theta0), float(sign))

    def abd(self):
        A = np.empty(9); B = np.empty(9); D = np.empty(9)
        self._L.orc_get_abd(self._p, _dp(A), _dp(B), _dp(D))
        return A.reshape(3, 3), B.reshape(3, 3), D.reshape(3, 3)

    def add_crease(self, vx, vy, ext_s=True, ext_e=True, closed=False):
        vx, vy = _f64(vx), _f64(vy)
        self._L.orc_add_crease(self._p, vx.size, _dp(vx), _dp(vy), int(ext_s), int(ext_e), int(closed))

    def add_domain(self, gps, is_face=False) -> int:
        a = [_f64(x) for x in (gps.gx, gps.gy, gps.w, gps.nx, gps.ny, gps.tx, gps.ty)]
        d = self._L.orc_add_domain(self._p, a[0].size, *[_dp(x) for x in a], int(is_face))
        self.domains.append(a[0].size)
        return d

    def add_fold(self, left, right, Kt, Ang0=0.0) -> int:
        return self._L.orc_add_fold(self._p, left, right, Kt, Ang0)

    def set_fold_gp_state(self, fold, theta0=None, active=None):
        """Per crease-GP rest angle and activity (None clears)."""
        l, _ = self.fold_doms[fold]
        n = self.domains[l]
        th = None if theta0 is None else _f64(theta0)
        ac = None if active is None else np.ascontiguousarray(active, dtype=np.int8)
        self._L.orc_set_fold_gp_state(self._p, fold, n, _dp(th),
                                      None if ac is None else ac.ctypes.data_as(_c))

    def set_fold_rigid(self, fold, rigid=True):
        self._L.orc_set_fold_rigid(self._p, fold, int(rigid))

    def indicator(self, c, x, y):
        out = np.empty(8)
        self._L.orc_indicator(self._p, c, float(x), float(y), _dp(out))
        return out

    # --- hot path --------------------------------------------------------
    def neighbours(self, dom, rho):
        npairs = self._L.orc_neighbours(self._p, dom, rho)
        ng = self.domains[dom]
        off = np.empty(ng + 1, np.int32); ids = np.empty(npairs, np.int32); s = np.empty(npairs)
        self._L.orc_get_neighbours(self._p, dom, _ip(off), _ip(ids), _dp(s))
        return off, ids, s

    def set_neighbours(self, dom, off, ids, s):
        off = np.ascontiguousarray(off, np.int32); ids = np.ascontiguousarray(ids, np.int32)
        s = _f64(s)
        self._L.orc_set_neighbours(self._p, dom, off.size - 1, _ip(off), _ip(ids), _dp(s))

    def shape_functions(self, dom, rho, order=2, kernel=0, nthreads=1):
        nb = C.c_int(0); fb = C.c_int(0)
        st = self._L.orc_shape_functions(self._p, dom, rho, order, kernel, nthreads, C.byref(nb), C.byref(fb))
        return st, nb.value, fb.value

    def get_shape_functions(self, dom, npairs):
        arrs = [np.empty(npairs) for _ in range(6)]
        self._L.orc_get_shape_functions(self._p, dom, *[_dp(a) for a in arrs])
        return arrs  # psi, dpx, dpy, dxx, dyy, dxy

    def init_assembly(self) -> int:
        return self._L.orc_init_assembly(self._p)

    def face_pattern(self):
        nnz = self._L.orc_face_nnz(self._p)
        outer = np.empty(3 * self.nn + 1, np.int32); inner = np.empty(nnz, np.int32)
        self._L.orc_get_face_pattern(self._p, _ip(outer), _ip(inner))
        return outer, inner

    def gp_map(self, g, n):
        out = np.empty(n * (n + 1) // 2, np.int32)
        self._L.orc_get_gp_map(self._p, g, _ip(out))
        return out

    def assemble(self, x, flag_rig=False, exact_tangent=False, nthreads=1):
        x = _f64(x)
        nnz = self._L.orc_assemble(self._p, _dp(x), int(flag_rig), int(exact_tangent), nthreads)
        outer = np.empty(3 * self.nn + 1, np.int32); inner = np.empty(nnz, np.int32)
        val = np.empty(nnz); F = np.empty(3 * self.nn)
        self._L.orc_get_system(self._p, _ip(outer), _ip(inner), _dp(val), _dp(F))
        return outer, inner, val, F

    def internal_force(self, x, flag_rig=False, with_tangent=True, nthreads=1):
        x = _f64(x)
        F = np.empty(3 * self.nn)
        self._L.orc_internal_force(self._p, _dp(x), int(flag_rig), int(with_tangent), nthreads, _dp(F))
        return F

    def last_seconds(self) -> float:
        return self._L.orc_last_seconds(self._p)

    def set_bc(self, dom, kind, active=0, val=(0, 0, 0, 0), force=(0, 0, 0), moment=0.0):
        """BoundaryCondition of a domain; kind 1 = point, 2 = edge, 0 = other."""
        v = _f64(np.asarray(val, float)); f = _f64(np.asarray(force, float))
        self._L.orc_set_domain_bc(self._p, dom, kind, int(active), _dp(v), _dp(f), float(moment))

    def constraints(self, x, flag_rig=False):
        """Meshless::evaluateConstraints: c and the COO triplets (rows, cols, vals)
        in push order."""
        x = _f64(x)
        nr = C.c_long(0)
        nt = self._L.orc_constraints(self._p, _dp(x), int(flag_rig), C.byref(nr))
        c = np.empty(nr.value); r = np.empty(nt, np.int32); col = np.empty(nt, np.int32); v = np.empty(nt)
        self._L.orc_get_constraints(self._p, _dp(c), _ip(r), _ip(col), _dp(v))
        return c, r, col, v

    def external_force(self):
        F = np.empty(3 * self.nn)
        self._L.orc_external_force(self._p, _dp(F))
        return F

    def energies(self, x):
        """(E_mem, E_ben, E_hin) at x."""
        x = _f64(x)
        out = np.empty(3)
        self._L.orc_energies(self._p, _dp(x), _dp(out))
        return tuple(out)

    def fold_angles(self, fold, x):
        x = _f64(x)
        l, _ = self.fold_doms[fold]
        th = np.empty(self.domains[l])
        self._L.orc_fold_angles(self._p, fold, _dp(x), _dp(th))
        return th

    def clear_creases(self):
        self._L.orc_clear_creases(self._p)

    def kinematics(self, dom, x):
        x = _f64(x)
        out = np.empty((self.domains[dom], 15))
        self._L.orc_kinematics(self._p, dom, _dp(x), _dp(out))
        return out

    def export_columns(self, gx, gy, x, rho, codes=()):
        """Exporter::writeBinaryStep's face-GP columns (plate/Exporter.h:47-91):
        X, Y, u, v, w (Psi-weighted sums of the displacements, in neighbour
        order), rotX = atan2(-b_y, b_z), rotY = atan2(b_x, b_z) in degrees, and
        E0 / E1 components for codes 0..5 (zeros otherwise)."""
        x = _f64(x)
        dom = self.face_dom
        off, ids, _ = self.neighbours(dom, rho)
        psi = self.get_shape_functions(dom, int(off[-1]))[0]
        kin = self.kinematics(dom, x)
        ng = off.size - 1
        out = np.zeros((ng, 7 + len(codes)))
        out[:, 0], out[:, 1] = gx, gy
        X = x.reshape(-1, 3)
        for g in range(ng):
            u = np.zeros(3)
            for j in range(off[g], off[g + 1]):
                u += X[ids[j]] * psi[j]
            out[g, 2:5] = u
        b = kin[:, 6:9]
        out[:, 5] = np.degrees(np.arctan2(-b[:, 1], b[:, 2]))
        out[:, 6] = np.degrees(np.arctan2(b[:, 0], b[:, 2]))
        for c, code in enumerate(codes):
            if 0 <= code <= 5:
                out[:, 7 + c] = kin[:, 9 + code]
        return out


def plate_from_sheet(sh, nthreads: int = 1, compute: bool = True):
    """Build the oracle plate for a workloads.Sheet: nodes, material, creases,
    face domain (0), fold domains (left, right per fold); optionally run
    FRNN + RKPM on every domain and initSparsityPattern."""
    p = OraclePlate()
    p.set_nodes(sh.X, sh.Y)
    p.set_plies(sh.plies)
    if getattr(sh, "natural_curvature", None) is not None:
        p.set_natural_curvature(True, *sh.natural_curvature)
    for cr in sh.creases:
        p.add_crease(cr.vx, cr.vy, cr.ext_s, cr.ext_e, cr.closed)
    face = p.add_domain(sh.face, is_face=True)
    fold_doms = []
    for k, f in enumerate(sh.folds):
        l = p.add_domain(f.left); r = p.add_domain(f.right)
        p.add_fold(l, r, f.Kt, f.Ang0)
        fold_doms.append((l, r))
        if getattr(f, "rigid", False):
            p.set_fold_rigid(k, True)
        for d, fv in ((l, getattr(f, "force_left", None)), (r, getattr(f, "force_right", None))):
            if fv is not None and any(v != 0.0 for v in fv):
                p.set_bc(d, 0, force=fv)
    p.fold_doms = fold_doms
    p.face_dom = face
    p.bnd_doms = []
    for bd in getattr(sh, "boundaries", []):
        d = p.add_domain(bd.gps)
        p.set_bc(d, 1 if bd.kind == "point" else 2, bd.active, bd.val, bd.force, bd.moment)
        p.bnd_doms.append(d)
    if compute:
        for d in range(len(p.domains)):
            p.neighbours(d, sh.rho)
            st, nb, fb = p.shape_functions(d, sh.rho, sh.order, sh.kernel, nthreads)
            if st != 0:
                raise RuntimeError(f"oracle RKPM status {st} on domain {d}: {nb} bad, first {fb}")
        st = p.init_assembly()
        if st != 0:
            raise RuntimeError("oracle pattern map miss")
    return p
