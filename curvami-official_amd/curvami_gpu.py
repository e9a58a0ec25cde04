"""ctypes binding of libcurvami_gpu.so (the C ABI declared in include/curvami_gpu.h).

`Plate` mirrors the reference's Meshless + NonLinearAssembler surface
(plate/Meshless.h:11-127, plate/Assembly.h:18-54): domains, folds, neighbour
search, shape functions, pattern, assembleSystem. Errors raise RuntimeError
with the library's message (the reference throws std::runtime_error with the
same text, e.g. core/include/curvami/GlobalRKPM.h:246-255).

There is no CPU fallback: if the library or a GPU is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcurvami_gpu.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)

# status codes (cvm_status)
OK, E_INVALID, E_HIP, E_SINGULAR_MOMENT, E_PIVOT_COLLAPSE, E_PATTERN_MISS, E_NO_MEMORY, E_STATE, E_BUDGET = range(9)
FACE, EDGE, POINT, FOLD_LEFT, FOLD_RIGHT = range(5)
RESIDUAL, TANGENT, ELASTIC_FOLDS, DEVICE_PTRS, EXACT_HINGE, FP32_TANGENT = 1, 2, 4, 8, 16, 32

# every exported symbol of include/curvami_gpu.h with its signature
SIGNATURES = {
    "cvm_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cvm_destroy": (None, [C.c_void_p]),
    "cvm_last_error": (C.c_char_p, [C.c_void_p]),
    "cvm_set_nodes": (C.c_int, [C.c_void_p, C.c_int, _d, _d]),
    "cvm_set_laminate": (C.c_int, [C.c_void_p, C.c_int, _d]),
    "cvm_set_natural_curvature": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "cvm_add_sheet": (C.c_int, [C.c_void_p, C.c_int, _d, _d, C.c_double, _i, _i]),
    "cvm_get_abd": (C.c_int, [C.c_void_p, _d, _d, _d]),
    "cvm_add_crease": (C.c_int, [C.c_void_p, C.c_int, _d, _d, C.c_int, C.c_int, C.c_int]),
    "cvm_add_domain": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _d, _d, _d, _d, _d, _d, _d, _i]),
    "cvm_add_fold": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, _i]),
    "cvm_set_fold_gp_state": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _d, C.c_char_p]),
    "cvm_set_fold_rigid": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "cvm_neighbours": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _i64]),
    "cvm_get_neighbours": (C.c_int, [C.c_void_p, C.c_int, _i, _i, _d]),
    "cvm_refresh_neighbours": (C.c_int, [C.c_void_p, _i]),
    "cvm_shape_functions": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _i, _i]),
    "cvm_get_shape_functions": (C.c_int, [C.c_void_p, C.c_int, _d, _d, _d, _d, _d, _d]),
    "cvm_shape_functions_all": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _i, _i, _i]),
    "cvm_recompute_shape_functions": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _i, _i, _i, _i]),
    "cvm_init_assembly": (C.c_int, [C.c_void_p, _i64, _i64]),
    "cvm_get_pattern": (C.c_int, [C.c_void_p, C.c_int, _i, _i]),
    "cvm_csr_mapping": (C.c_int, [C.c_void_p, C.c_int, _i, C.c_int, _i]),
    "cvm_assemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "cvm_assemble_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "cvm_set_domain_bc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _d, _d, C.c_double]),
    "cvm_constraint_layout": (C.c_int, [C.c_void_p, C.c_int, _i64, _i64]),
    "cvm_get_constraint_pattern": (C.c_int, [C.c_void_p, _i, _i]),
    "cvm_constraints": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "cvm_external_force": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cvm_energies": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "cvm_fold_angles": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "cvm_clear_creases": (C.c_int, [C.c_void_p]),
    "cvm_export_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _i, C.c_void_p]),
    "cvm_restart_last_error": (C.c_char_p, []),
    "cvm_restart_save": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                   _d, _d]),
    "cvm_restart_load": (C.c_int, [C.c_char_p, C.c_int, _i, _d, _i, _d, _d, C.c_int]),
    "cvm_synchronize": (C.c_int, [C.c_void_p]),
    "cvm_set_tile_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "cvm_get_tile_shape": (C.c_int, [C.c_void_p, _i]),
    "cvm_graph_capture": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cvm_graph_launch": (C.c_int, [C.c_void_p]),
    "cvm_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_void_p)]),
    "cvm_server_destroy": (None, [C.c_void_p]),
    "cvm_server_last_error": (C.c_char_p, [C.c_void_p]),
    "cvm_server_add_sheet": (C.c_int, [C.c_void_p, C.c_void_p, _i]),
    "cvm_server_request_bytes": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _i64]),
    "cvm_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _i64]),
    "cvm_server_poll": (C.c_int, [C.c_void_p, C.c_int64, _i]),
    "cvm_server_wait": (C.c_int, [C.c_void_p, C.c_int64, _d]),
    "cvm_server_stats": (C.c_int, [C.c_void_p, _i64]),
    "cvm_server_set_batching": (C.c_int, [C.c_void_p, C.c_int]),
    "cvm_server_batching_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "cvm_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cvm_kernel_count": (C.c_int, [C.c_void_p]),
    "cvm_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, _d, _i64]),
    "cvm_reset_profiling": (C.c_int, [C.c_void_p]),
    "cvm_mesh_last_error": (C.c_char_p, []),
    "cvm_mesh_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "cvm_mesh_destroy": (None, [C.c_void_p]),
    "cvm_mesh_info": (C.c_int, [C.c_void_p, _i64]),
    "cvm_mesh_nodes": (C.c_int, [C.c_void_p, _d, _d]),
    "cvm_mesh_elements": (C.c_int, [C.c_void_p, _i, _i, _i64, _i]),
    "cvm_mesh_quadrature_sizes": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _i64]),
    "cvm_mesh_quadrature": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _i, _d, _i, _d, _d, _i, _d, _i]),
    "cvm_mesh_extract_curve": (C.c_int, [C.c_void_p, _i, C.c_int, C.c_int64, _d, _d, _i64, _i, _i64]),
    "cvm_support_radius": (C.c_int, [_d, _d, C.c_int64, C.c_int, _d]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback for the hot path)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class CvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _dp(a):
    return None if a is None else a.ctypes.data_as(_d)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Plate:
    """Device-resident creased plate: nodes, laminate, creases, domains, folds."""

    def __init__(self, device: int = 0):
        self._L = load_library()
        h = C.c_void_p()
        st = self._L.cvm_create(device, C.byref(h))
        if st != OK:
            raise CvmError(st, f"cvm_create failed with status {st} (no HIP device?)")
        self._h = h
        self.nn = 0
        self.domain_sizes = []
        self.n_crows = self.n_ctrip = 0
        self.npairs = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.cvm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise CvmError(st, self._L.cvm_last_error(self._h).decode(errors="replace"))

    # ---- setup ----------------------------------------------------------------
    def set_nodes(self, X, Y):
        X, Y = _f64(X), _f64(Y)
        self.nn = X.size
        self._check(self._L.cvm_set_nodes(self._h, X.size, _dp(X), _dp(Y)))

    def set_plies(self, plies):
        plies = _f64(plies)
        self._check(self._L.cvm_set_laminate(self._h, plies.shape[0], _dp(plies)))

    def set_natural_curvature(self, active=True, x0=0.0, width=1.0, theta0=0.0, sign=-1.0):
        """Kinematics::naturalCurvature() (plate/Kinematics.h:20-45): the band
        sign * theta0 * g(X) in E1_el(0, 0); sign -1 is a mountain, +1 a valley."""
        self._check(self._L.cvm_set_natural_curvature(self._h, int(active), float(x0), float(width), float(theta0),
                                                      float(sign)))

    def abd(self):
        A, B, D = np.empty(9), np.empty(9), np.empty(9)
        self._check(self._L.cvm_get_abd(self._h, _dp(A), _dp(B), _dp(D)))
        return A.reshape(3, 3), B.reshape(3, 3), D.reshape(3, 3)

    def add_crease(self, vx, vy, ext_s=True, ext_e=True, closed=False):
        vx, vy = _f64(vx), _f64(vy)
        self._check(self._L.cvm_add_crease(self._h, vx.size, _dp(vx), _dp(vy), int(ext_s), int(ext_e), int(closed)))

    def clear_creases(self):
        self._check(self._L.cvm_clear_creases(self._h))

    def add_domain(self, gps, kind=FACE) -> int:
        a = [_f64(x) for x in (gps.gx, gps.gy, gps.w, gps.nx, gps.ny, gps.tx, gps.ty)]
        did = C.c_int(-1)
        self._check(self._L.cvm_add_domain(self._h, kind, a[0].size, *[_dp(x) for x in a], C.byref(did)))
        if did.value < len(self.domain_sizes):       # batched: face GPs appended to the face domain
            self.domain_sizes[did.value] += a[0].size
        else:
            self.domain_sizes.append(a[0].size)
        return did.value

    def add_sheet(self, X, Y, rho):
        """Batched context: append a sheet (nodes, support radius); returns
        (sheet id, id of its first node)."""
        X, Y = _f64(X), _f64(Y)
        sid = C.c_int(-1); n0 = C.c_int(-1)
        self._check(self._L.cvm_add_sheet(self._h, X.size, _dp(X), _dp(Y), float(rho), C.byref(sid), C.byref(n0)))
        self.nn += X.size
        return sid.value, n0.value

    def add_fold(self, left, right, Kt, Ang0=0.0) -> int:
        fid = C.c_int(-1)
        self._check(self._L.cvm_add_fold(self._h, left, right, Kt, Ang0, C.byref(fid)))
        return fid.value

    def set_fold_rigid(self, fold, rigid=True):
        self._check(self._L.cvm_set_fold_rigid(self._h, fold, int(rigid)))

    def set_fold_gp_state(self, fold, theta0=None, active=None):
        n = self.domain_sizes[0] if theta0 is None and active is None else \
            (len(theta0) if theta0 is not None else len(active))
        th = _f64(theta0)
        ac = None if active is None else np.ascontiguousarray(active, dtype=np.int8).tobytes()
        self._check(self._L.cvm_set_fold_gp_state(self._h, fold, n, _dp(th), ac))

    # ---- hot path -------------------------------------------------------------
    def neighbours(self, dom, rho):
        n = C.c_int64(0)
        self._check(self._L.cvm_neighbours(self._h, dom, rho, C.byref(n)))
        self.npairs[dom] = n.value
        return n.value

    def refresh_neighbours(self):
        """The neighbour search of every searched domain again on the device;
        returns True if any list changed (then the shape functions and, for
        face and fold domains, the assembly plan must be rebuilt)."""
        ch = C.c_int(0)
        self._check(self._L.cvm_refresh_neighbours(self._h, C.byref(ch)))
        if ch.value:
            for d in range(len(self.domain_sizes)):
                if d in self.npairs:
                    off = np.empty(self.domain_sizes[d] + 1, np.int32)
                    self._check(self._L.cvm_get_neighbours(self._h, d, off.ctypes.data_as(_i), None, None))
                    self.npairs[d] = int(off[-1])
        return bool(ch.value)

    def get_neighbours(self, dom):
        ng = self.domain_sizes[dom]
        npairs = self.npairs[dom]
        off = np.empty(ng + 1, np.int32); ids = np.empty(max(npairs, 1), np.int32); s = np.empty(max(npairs, 1))
        self._check(self._L.cvm_get_neighbours(self._h, dom, off.ctypes.data_as(_i), ids.ctypes.data_as(_i), _dp(s)))
        return off, ids[:npairs], s[:npairs]

    def shape_functions(self, dom, order=2, kernel=0):
        nb, fb = C.c_int(0), C.c_int(0)
        st = self._L.cvm_shape_functions(self._h, dom, order, kernel, C.byref(nb), C.byref(fb))
        if st in (E_SINGULAR_MOMENT, E_PIVOT_COLLAPSE):
            err = CvmError(st, self._L.cvm_last_error(self._h).decode())
            err.nbad, err.first = nb.value, fb.value
            raise err
        self._check(st)

    def shape_functions_all(self, order=2, kernel=0):
        """Every domain, one status read-back; raises with .domain, .nbad, .first."""
        dm, nb, fb = C.c_int(-1), C.c_int(0), C.c_int(0)
        st = self._L.cvm_shape_functions_all(self._h, order, kernel, C.byref(dm), C.byref(nb), C.byref(fb))
        if st in (E_SINGULAR_MOMENT, E_PIVOT_COLLAPSE):
            err = CvmError(st, self._L.cvm_last_error(self._h).decode())
            err.domain, err.nbad, err.first = dm.value, nb.value, fb.value
            raise err
        self._check(st)

    def recompute_shape_functions(self, order=2, kernel=0):
        """The neighbour search and the shape functions of every domain in one
        call (refresh_neighbours + shape_functions_all); returns True if any
        list changed, raises as shape_functions_all."""
        ch, dm, nb, fb = C.c_int(0), C.c_int(-1), C.c_int(0), C.c_int(0)
        st = self._L.cvm_recompute_shape_functions(self._h, order, kernel, C.byref(ch), C.byref(dm), C.byref(nb),
                                                   C.byref(fb))
        if st in (E_SINGULAR_MOMENT, E_PIVOT_COLLAPSE):
            err = CvmError(st, self._L.cvm_last_error(self._h).decode())
            err.domain, err.nbad, err.first = dm.value, nb.value, fb.value
            raise err
        self._check(st)
        if ch.value:
            for d in range(len(self.domain_sizes)):
                if d in self.npairs:
                    off = np.empty(self.domain_sizes[d] + 1, np.int32)
                    self._check(self._L.cvm_get_neighbours(self._h, d, off.ctypes.data_as(_i), None, None))
                    self.npairs[d] = int(off[-1])
        return bool(ch.value)

    def get_shape_functions(self, dom):
        n = max(self.npairs[dom], 1)
        arrs = [np.empty(n) for _ in range(6)]
        self._check(self._L.cvm_get_shape_functions(self._h, dom, *[_dp(a) for a in arrs]))
        return [a[:self.npairs[dom]] for a in arrs]

    def init_assembly(self):
        nf, nu = C.c_int64(0), C.c_int64(0)
        self._check(self._L.cvm_init_assembly(self._h, C.byref(nf), C.byref(nu)))
        self.nnz_face, self.nnz_union = nf.value, nu.value
        return self.nnz_face, self.nnz_union

    def pattern(self, with_folds=True):
        nnz = self.nnz_union if with_folds else self.nnz_face
        outer = np.empty(3 * self.nn + 1, np.int32); inner = np.empty(max(nnz, 1), np.int32)
        self._check(self._L.cvm_get_pattern(self._h, int(with_folds), outer.ctypes.data_as(_i), inner.ctypes.data_as(_i)))
        return outer, inner[:nnz]

    def csr_mapping(self, gp):
        n = C.c_int(0)
        self._check(self._L.cvm_csr_mapping(self._h, gp, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.int32)
        self._check(self._L.cvm_csr_mapping(self._h, gp, out.ctypes.data_as(_i), n.value, C.byref(n)))
        return out[:n.value]

    def assemble(self, x, elastic_folds=True, tangent=True, exact_hinge=False, fp32_tangent=False):
        """Meshless::assembleSystem: returns (F_int, K_values) in the pattern
        returned by pattern(with_folds=elastic_folds)."""
        x = _f64(x)
        F = np.empty(3 * self.nn)
        nnz = self.nnz_union if elastic_folds else self.nnz_face
        K = np.empty(max(nnz, 1)) if tangent else None
        flags = RESIDUAL | (TANGENT if tangent else 0) | (ELASTIC_FOLDS if elastic_folds else 0) \
            | (EXACT_HINGE if exact_hinge else 0) | (FP32_TANGENT if fp32_tangent else 0)
        self._check(self._L.cvm_assemble(self._h, x.ctypes.data, flags, F.ctypes.data,
                                         K.ctypes.data if K is not None else None))
        return F, (K[:nnz] if tangent else None)

    def assemble_device(self, x_ptr, F_ptr, K_ptr, flags):
        """Device-pointer variant (inputs already resident in HBM)."""
        self._check(self._L.cvm_assemble(self._h, C.c_void_p(x_ptr), flags | DEVICE_PTRS,
                                         C.c_void_p(F_ptr), C.c_void_p(K_ptr) if K_ptr else None))

    def assemble_noalloc(self, x_ptr, F_ptr, K_ptr, flags):
        """cvm_assemble_noalloc: device pointers, no device allocation (the
        library's static scratch; CVM_E_BUDGET when the sheet needs more)."""
        self._check(self._L.cvm_assemble_noalloc(self._h, C.c_void_p(x_ptr), flags, C.c_void_p(F_ptr),
                                                 C.c_void_p(K_ptr) if K_ptr else None))

    # ---- constraints and loads ----------------------------------------------
    def set_bc(self, dom, active=0, val=(0.0, 0.0, 0.0, 0.0), force=(0.0, 0.0, 0.0), moment=0.0):
        """BoundaryCondition of a domain (active: bit 0 u, 1 v, 2 w, 3 rotation)."""
        v = np.ascontiguousarray(val, np.float64); f = np.ascontiguousarray(force, np.float64)
        self._check(self._L.cvm_set_domain_bc(self._h, dom, int(active), v.ctypes.data_as(_d),
                                              f.ctypes.data_as(_d), float(moment)))

    def constraint_layout(self, rigid=False):
        """Builds the constraint-row layout; returns the triplet (rows, cols)."""
        nr = C.c_int64(0); nt = C.c_int64(0)
        self._check(self._L.cvm_constraint_layout(self._h, int(rigid), C.byref(nr), C.byref(nt)))
        self.n_crows, self.n_ctrip = nr.value, nt.value
        rows = np.empty(max(nt.value, 1), np.int32); cols = np.empty(max(nt.value, 1), np.int32)
        self._check(self._L.cvm_get_constraint_pattern(self._h, rows.ctypes.data_as(_i), cols.ctypes.data_as(_i)))
        return rows[:nt.value], cols[:nt.value]

    def constraints(self, x):
        """c(x) and the triplet values at x (layout from constraint_layout)."""
        x = _f64(x)
        c = np.empty(max(self.n_crows, 1)); v = np.empty(max(self.n_ctrip, 1))
        self._check(self._L.cvm_constraints(self._h, x.ctypes.data, 0, c.ctypes.data, v.ctypes.data))
        return c[:self.n_crows], v[:self.n_ctrip]

    def external_force(self):
        F = np.empty(3 * self.nn)
        self._check(self._L.cvm_external_force(self._h, 0, F.ctypes.data))
        return F

    def energies(self, x):
        """(E_mem, E_ben, E_hin) at x (Meshless::computeStrainEnergies,
        computeTotalHingeEnergy)."""
        x = _f64(x)
        out = np.empty(3)
        self._check(self._L.cvm_energies(self._h, x.ctypes.data, 0, out.ctypes.data))
        return tuple(float(v) for v in out)

    def export_fields(self, x, codes=()):
        """(face GPs, 7 + len(codes)) exporter columns: X, Y, u, v, w, rotX, rotY
        (degrees) and one column per field code (Exporter::writeBinaryStep)."""
        x = _f64(x)
        codes = np.ascontiguousarray(codes, dtype=np.int32)
        ng = self.domain_sizes[self.face_dom]
        out = np.empty((max(ng, 1), 7 + codes.size))
        self._check(self._L.cvm_export_fields(self._h, x.ctypes.data, 0, int(codes.size),
                                              codes.ctypes.data_as(_i) if codes.size else None, out.ctypes.data))
        return out[:ng]

    def fold_angles(self, fold, x):
        """Dihedral angle at every crease GP of a fold."""
        x = _f64(x)
        l, _ = self.fold_doms[fold]
        th = np.empty(max(self.domain_sizes[l], 1))
        self._check(self._L.cvm_fold_angles(self._h, int(fold), x.ctypes.data, 0, th.ctypes.data))
        return th[:self.domain_sizes[l]]

    def synchronize(self):
        self._check(self._L.cvm_synchronize(self._h))

    def set_tile_shape(self, tile_gps=16, union_max=0, grid=0):
        self._check(self._L.cvm_set_tile_shape(self._h, int(tile_gps), int(union_max), int(grid)))

    def tile_shape(self):
        out = (C.c_int * 4)()
        self._check(self._L.cvm_get_tile_shape(self._h, out))
        return {"tile_gps": out[0], "union_max": out[1], "grid": out[2], "ntiles": out[3]}

    def graph_capture(self, x_ptr, F_ptr, K_ptr, flags):
        self._check(self._L.cvm_graph_capture(self._h, flags, C.c_void_p(x_ptr), C.c_void_p(F_ptr),
                                              C.c_void_p(K_ptr) if K_ptr else None))

    def graph_launch(self):
        self._check(self._L.cvm_graph_launch(self._h))

    # ---- measurement ----------------------------------------------------------
    def set_profiling(self, on=True):
        self._check(self._L.cvm_set_profiling(self._h, int(on)))

    def reset_profiling(self):
        self._check(self._L.cvm_reset_profiling(self._h))

    def kernel_stats(self):
        out = {}
        for i in range(self._L.cvm_kernel_count(self._h)):
            name = C.create_string_buffer(128)
            ms = C.c_double(0); cnt = C.c_int64(0)
            self._check(self._L.cvm_kernel_stats(self._h, i, name, 128, C.byref(ms), C.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out


def batch_from_sheets(sheets, device: int = 0, compute: bool = True):
    """One batched context holding every workloads.Sheet (same laminate, order
    and window for all): per-sheet nodes, creases, face GPs (shared face
    domain), fold domains and boundary domains. Attributes: node0 (first node
    id per sheet), face_range ((first GP, count) per sheet in the face domain)."""
    p = Plate(device)
    p.set_plies(sheets[0].plies)
    p.node0, p.face_range, p.fold_doms = [], [], []
    face = None
    for sh in sheets:
        _, n0 = p.add_sheet(sh.X, sh.Y, sh.rho)
        p.node0.append(n0)
        for cr in sh.creases:
            p.add_crease(cr.vx, cr.vy, cr.ext_s, cr.ext_e, cr.closed)
        before = p.domain_sizes[face] if face is not None else 0
        face = p.add_domain(sh.face, FACE)
        p.face_range.append((before, sh.face.n))
        for f in sh.folds:
            l = p.add_domain(f.left, FOLD_LEFT)
            r = p.add_domain(f.right, FOLD_RIGHT)
            p.add_fold(l, r, f.Kt, f.Ang0)
            p.fold_doms.append((l, r))
    p.face_dom = face
    if compute:
        for d in range(len(p.domain_sizes)):
            p.neighbours(d, 0.0)
        p.shape_functions_all(sheets[0].order, sheets[0].kernel)
        p.init_assembly()
    return p


def plate_from_sheet(sh, device: int = 0, compute: bool = True, order=None, kernel=None):
    """Build the device plate for a workloads.Sheet: face domain 0, fold domains
    (left, right) per fold; optionally neighbours + shape functions on every
    domain and the assembly plan."""
    p = Plate(device)
    p.set_nodes(sh.X, sh.Y)
    p.set_plies(sh.plies)
    if getattr(sh, "natural_curvature", None) is not None:
        p.set_natural_curvature(True, *sh.natural_curvature)
    for cr in sh.creases:
        p.add_crease(cr.vx, cr.vy, cr.ext_s, cr.ext_e, cr.closed)
    p.face_dom = p.add_domain(sh.face, FACE)
    p.fold_doms = []
    for k, f in enumerate(sh.folds):
        l = p.add_domain(f.left, FOLD_LEFT)
        r = p.add_domain(f.right, FOLD_RIGHT)
        p.add_fold(l, r, f.Kt, f.Ang0)
        p.fold_doms.append((l, r))
        if getattr(f, "rigid", False):
            p.set_fold_rigid(k, True)
        for d, fv in ((l, getattr(f, "force_left", None)), (r, getattr(f, "force_right", None))):
            if fv is not None and any(v != 0.0 for v in fv):
                p.set_bc(d, 0, force=fv)
    p.bnd_doms = []
    for bd in getattr(sh, "boundaries", []):
        d = p.add_domain(bd.gps, POINT if bd.kind == "point" else EDGE)
        p.set_bc(d, bd.active, bd.val, bd.force, bd.moment)
        p.bnd_doms.append(d)
    if compute:
        for d in range(len(p.domain_sizes)):
            p.neighbours(d, sh.rho)
            p.shape_functions(d, sh.order if order is None else order, sh.kernel if kernel is None else kernel)
        p.init_assembly()
    return p


def _mesh_check(L, st):
    if st != OK:
        raise CvmError(st, L.cvm_mesh_last_error().decode(errors="replace"))


def _ip(a):
    return a.ctypes.data_as(_i)


class MeshQuadrature:
    """Gauss points of a mesh (Quadrature::generate, core/include/curvami/Quadrature.h:59-233):
    active = raw indices of the face-element nodes in ascending order (node i of
    the sheet is raw node active[i]); face/edge/point arrays hold x, y, w per
    GP, edge_tn holds tx, ty, nx, ny; *_tag the element's entity tag."""

    def __init__(self, active, face, face_tag, edge, edge_tn, edge_tag, point, point_tag):
        self.active, self.face, self.face_tag = active, face, face_tag
        self.edge, self.edge_tn, self.edge_tag = edge, edge_tn, edge_tag
        self.point, self.point_tag = point, point_tag


class Mesh:
    """A Gmsh 4.1 ASCII mesh read by the library (GmshReader::readMesh,
    core/include/curvami/GmshReader.h:32-146). Host-only: no GPU is needed."""

    def __init__(self, path: str):
        self._L = load_library()
        h = C.c_void_p()
        _mesh_check(self._L, self._L.cvm_mesh_read(os.fsencode(path), C.byref(h)))
        self._h = h
        self.path = path
        info = np.zeros(3, np.int64)
        _mesh_check(self._L, self._L.cvm_mesh_info(h, info.ctypes.data_as(_i64)))
        nn, nel, nent = (int(v) for v in info)
        self.X = np.empty(nn); self.Y = np.empty(nn)
        _mesh_check(self._L, self._L.cvm_mesh_nodes(h, _dp(self.X), _dp(self.Y)))
        self.type = np.empty(nel, np.int32); self.entity = np.empty(nel, np.int32)
        self.node_off = np.empty(nel + 1, np.int64); self.node_idx = np.empty(max(nent, 1), np.int32)
        _mesh_check(self._L, self._L.cvm_mesh_elements(h, _ip(self.type), _ip(self.entity),
                                                       self.node_off.ctypes.data_as(_i64), _ip(self.node_idx)))
        self.node_idx = self.node_idx[:nent]

    def close(self):
        if getattr(self, "_h", None):
            self._L.cvm_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def quadrature(self, line_degree: int = 1, face_degree: int = 3) -> MeshQuadrature:
        sz = np.zeros(4, np.int64)
        _mesh_check(self._L, self._L.cvm_mesh_quadrature_sizes(self._h, line_degree, face_degree,
                                                               sz.ctypes.data_as(_i64)))
        na, nf, ne, npt = (int(v) for v in sz)
        q = MeshQuadrature(np.empty(max(na, 1), np.int32), np.empty((max(nf, 1), 3)), np.empty(max(nf, 1), np.int32),
                           np.empty((max(ne, 1), 3)), np.empty((max(ne, 1), 4)), np.empty(max(ne, 1), np.int32),
                           np.empty((max(npt, 1), 3)), np.empty(max(npt, 1), np.int32))
        _mesh_check(self._L, self._L.cvm_mesh_quadrature(
            self._h, line_degree, face_degree, _ip(q.active), _dp(q.face), _ip(q.face_tag), _dp(q.edge),
            _dp(q.edge_tn), _ip(q.edge_tag), _dp(q.point), _ip(q.point_tag)))
        q.active = q.active[:na]
        q.face, q.face_tag = q.face[:nf], q.face_tag[:nf]
        q.edge, q.edge_tn, q.edge_tag = q.edge[:ne], q.edge_tn[:ne], q.edge_tag[:ne]
        q.point, q.point_tag = q.point[:npt], q.point_tag[:npt]
        return q

    def extract_curve(self, tags):
        """Crease polyline of the line elements with these entity tags
        (MeshUtils::extractCurve, core/include/curvami/MeshUtils.h:57-134):
        (vx, vy, closed, extend_start, extend_end, dropped)."""
        t = np.ascontiguousarray(np.atleast_1d(tags), dtype=np.int32)
        nv = C.c_int64(); fl = C.c_int(); dr = C.c_int64()
        _mesh_check(self._L, self._L.cvm_mesh_extract_curve(self._h, _ip(t), t.size, 0, None, None,
                                                            C.byref(nv), C.byref(fl), C.byref(dr)))
        vx = np.empty(max(nv.value, 1)); vy = np.empty(max(nv.value, 1))
        _mesh_check(self._L, self._L.cvm_mesh_extract_curve(self._h, _ip(t), t.size, nv.value, _dp(vx), _dp(vy),
                                                            C.byref(nv), C.byref(fl), C.byref(dr)))
        n = nv.value
        return vx[:n], vy[:n], bool(fl.value & 1), bool(fl.value & 2), bool(fl.value & 4), int(dr.value)


def support_radius(X, Y, order: int) -> float:
    """(order + 3) x median 3rd-nearest-neighbour spacing
    (MeshUtils::computeSupportRadius, core/include/curvami/MeshUtils.h:144-209)."""
    L = load_library()
    X = _f64(X); Y = _f64(Y)
    if X.shape != Y.shape:
        raise ValueError("X and Y differ in size")
    r = C.c_double()
    _mesh_check(L, L.cvm_support_radius(_dp(X), _dp(Y), X.size, int(order), C.byref(r)))
    return r.value


def restart_save(path, k, lam, ds, dl_old, W_ext, x, dU_old):
    """RestartManager::save: the binary .res record."""
    L = load_library()
    x = _f64(x); dU_old = _f64(dU_old)
    if x.shape != dU_old.shape:
        raise CvmError(E_INVALID, "Restart save refused: x and dU_old differ in size.")
    st = L.cvm_restart_save(os.fsencode(path), int(k), float(lam), float(ds), float(dl_old), float(W_ext),
                            int(x.size), _dp(x), _dp(dU_old))
    if st != OK:
        raise CvmError(st, L.cvm_restart_last_error().decode(errors="replace"))


def restart_load(path, expected_ns: int = -1):
    """RestartManager::load: dict(k, lam, ds, dl_old, W_ext, x, dU_old); raises
    CvmError for a foreign, truncated, mismatched or non-finite record."""
    L = load_library()
    k = C.c_int(); ns = C.c_int(); sc = np.zeros(4)
    st = L.cvm_restart_load(os.fsencode(path), int(expected_ns), C.byref(k), _dp(sc), C.byref(ns), None, None, 0)
    if st != OK:
        raise CvmError(st, L.cvm_restart_last_error().decode(errors="replace"))
    x = np.empty(ns.value); dU = np.empty(ns.value)
    st = L.cvm_restart_load(os.fsencode(path), int(expected_ns), C.byref(k), _dp(sc), C.byref(ns), _dp(x), _dp(dU),
                            ns.value)
    if st != OK:
        raise CvmError(st, L.cvm_restart_last_error().decode(errors="replace"))
    return {"k": k.value, "lam": sc[0], "ds": sc[1], "dl_old": sc[2], "W_ext": sc[3], "x": x, "dU_old": dU}


class DeviceArray:
    """A float64 device buffer (hipMalloc through the HIP runtime); plumbing
    for device-pointer entry points and the serving loop."""
    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            H = C.CDLL("libamdhip64.so")
            H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            H.hipFree.argtypes = [C.c_void_p]
            H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = H
        return cls._hip

    def __init__(self, n: int):
        self.n = int(n)
        p = C.c_void_p()
        if self._rt().hipMalloc(C.byref(p), C.c_size_t(max(self.n, 1) * 8)) != 0:
            raise CvmError(E_NO_MEMORY, f"hipMalloc of {self.n * 8} bytes failed")
        self.ptr = p.value

    def get(self) -> np.ndarray:
        out = np.empty(self.n)
        if self.n and self._rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.n * 8, 2) != 0:
            raise CvmError(E_HIP, "hipMemcpy D2H failed")
        return out

    def get_range(self, lo: int, hi: int) -> np.ndarray:
        """Elements [lo, hi) copied to the host."""
        lo, hi = int(lo), int(hi)
        if not 0 <= lo <= hi <= self.n:
            raise ValueError("range outside the buffer")
        out = np.empty(hi - lo)
        if hi > lo and self._rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + 8 * lo),
                                            (hi - lo) * 8, 2) != 0:
            raise CvmError(E_HIP, "hipMemcpy D2H failed")
        return out

    def set(self, a):
        a = _f64(a)
        if a.size != self.n:
            raise ValueError("size mismatch")
        if self.n and self._rt().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.n * 8, 1) != 0:
            raise CvmError(E_HIP, "hipMemcpy H2D failed")

    def free(self):
        if getattr(self, "ptr", None):
            self._rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """The serving loop (cvm_server_*): prepared plates are registered as
    sheets; submit() queues an evaluation and returns a request id; wait()
    returns (status, queue_ms, device_ms, total_ms)."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 8, device: int = 0):
        self._L = load_library()
        h = C.c_void_p()
        st = self._L.cvm_server_create(device, int(hbm_budget_bytes), int(max_inflight), C.byref(h))
        if st != OK:
            raise CvmError(st, f"cvm_server_create failed with status {st}")
        self._h = h
        self._plates = []
        self._keep = {}

    def _check(self, st):
        if st != OK:
            raise CvmError(st, self._L.cvm_server_last_error(self._h).decode())

    def add_sheet(self, plate) -> int:
        sid = C.c_int(-1)
        self._check(self._L.cvm_server_add_sheet(self._h, plate._h, C.byref(sid)))
        self._plates.append(plate)
        return sid.value

    def request_bytes(self, sheet, flags, with_K=True) -> int:
        b = C.c_int64(0)
        self._check(self._L.cvm_server_request_bytes(self._h, sheet, flags, int(with_K), C.byref(b)))
        return b.value

    def submit(self, sheet, x, F_out, flags=RESIDUAL | TANGENT | ELASTIC_FOLDS, K_dev=None) -> int:
        x = _f64(x)
        rid = C.c_int64(0)
        self._check(self._L.cvm_server_submit(self._h, sheet, x.ctypes.data, flags, F_out.ctypes.data,
                                              C.c_void_p(K_dev.ptr) if K_dev is not None else None,
                                              C.byref(rid)))
        self._keep[rid.value] = (x, F_out, K_dev)
        return rid.value

    def wait(self, rid):
        t = (C.c_double * 3)()
        st = self._L.cvm_server_wait(self._h, C.c_int64(rid), t)
        self._keep.pop(rid, None)
        self._check(st)
        return t[0], t[1], t[2]

    def stats(self):
        out = (C.c_int64 * 4)()
        self._check(self._L.cvm_server_stats(self._h, out))
        return {"completed": out[0], "inflight": out[1], "pending": out[2], "workspace_bytes": out[3]}

    def set_batching(self, on=True):
        self._check(self._L.cvm_server_set_batching(self._h, int(on)))

    def batching_stats(self):
        out = (C.c_int64 * 2)()
        self._check(self._L.cvm_server_batching_stats(self._h, out))
        return {"shared_tile_launches": out[0], "requests_in_shared_launches": out[1]}

    def close(self):
        if getattr(self, "_h", None):
            self._L.cvm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Host KKT solver (libcurvami_kkt.so, include/curvami_kkt.h)
# ---------------------------------------------------------------------------
_KKT_PATH = os.path.join(_HERE, "libcurvami_kkt.so")
_kkt_lib = None
E_KKT = 9


def load_kkt_library(path: str = _KKT_PATH):
    global _kkt_lib
    if _kkt_lib is not None:
        return _kkt_lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "cvm_kkt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
        "cvm_kkt_destroy": (None, [P]),
        "cvm_kkt_last_error": (C.c_char_p, [P]),
        "cvm_kkt_set_pattern": (C.c_int, [P, C.c_int, _i, _i, C.c_int, C.c_int64, _i, _i]),
        "cvm_kkt_factorize": (C.c_int, [P, _d, _d, _i, _i]),
        "cvm_kkt_solve": (C.c_int, [P, C.c_int, _d, _d]),
        "cvm_kkt_stats": (C.c_int, [P, C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _kkt_lib = L
    return L


class KKT:
    """[K J^T; J 0] by a sparse symmetric indefinite LDL^T on the host
    (MumpsSymmetricSolver's role): set_pattern (re-analysis only when the
    pattern changes), factorize -> (negative, positive pivots), solve."""

    def __init__(self, nthreads: int = 0):
        self._L = load_kkt_library()
        h = C.c_void_p()
        st = self._L.cvm_kkt_create(int(nthreads), C.byref(h))
        if st != OK:
            raise CvmError(st, "cvm_kkt_create failed")
        self._h = h
        self.n = self.m = 0

    def _check(self, st):
        if st != OK:
            raise CvmError(st, self._L.cvm_kkt_last_error(self._h).decode(errors="replace"))

    def set_pattern(self, outer, inner, m, rows, cols):
        outer = np.ascontiguousarray(outer, np.int32); inner = np.ascontiguousarray(inner, np.int32)
        rows = np.ascontiguousarray(rows, np.int32); cols = np.ascontiguousarray(cols, np.int32)
        self.n, self.m = outer.size - 1, int(m)
        self._check(self._L.cvm_kkt_set_pattern(self._h, self.n, outer.ctypes.data_as(_i), inner.ctypes.data_as(_i),
                                                self.m, rows.size, rows.ctypes.data_as(_i), cols.ctypes.data_as(_i)))

    def factorize(self, K_values, J_values):
        K_values = _f64(K_values); J_values = _f64(J_values)
        neg, pos = C.c_int(0), C.c_int(0)
        self._check(self._L.cvm_kkt_factorize(self._h, _dp(K_values), _dp(J_values) if J_values.size else None,
                                              C.byref(neg), C.byref(pos)))
        return neg.value, pos.value

    def solve(self, rhs):
        rhs = np.asfortranarray(rhs, dtype=np.float64)
        nrhs = 1 if rhs.ndim == 1 else rhs.shape[1]
        x = np.empty_like(rhs, order="F")
        self._check(self._L.cvm_kkt_solve(self._h, nrhs, rhs.ctypes.data_as(_d), x.ctypes.data_as(_d)))
        return x

    def stats(self):
        out = (C.c_int64 * 10)()
        self._check(self._L.cvm_kkt_stats(self._h, out))
        keys = ("analyses", "factorizations", "solves", "kkt_entries", "factor_entries", "perturbed_pivots",
                "analysis_us", "factor_us", "lu_fallbacks", "probe_backward_error_digits")
        return dict(zip(keys, list(out)))

    def close(self):
        if getattr(self, "_h", None):
            self._L.cvm_kkt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
