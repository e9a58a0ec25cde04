"""Step output (Exporter::writeBinaryStep, plate/Exporter.h:27-339).

Per step: the face-GP table [X, Y, u, v, w, rotX_deg, rotY_deg, fields...]
(device: cvm_export_fields), the per-crease dihedral table [fold_index,
crease_tag, X_ref, Y_ref, s_arc, theta_deg, weight] (device angles,
cvm_fold_angles; crease GPs chained nearest-neighbour from one end as in
:164-201), and the step attributes (angles, is_driver, disp, W_ext, E_mem,
E_ben, E_hin, planar_rms, planar_max, ref_diag, git_commit; :93-129, :231-322).

StepWriter appends each step to an HDF5 file with the reference's layout
(`/step_NNNNN`, `/step_NNNNN_creases` and their attributes, chunked and
deflate-compressed) through the library's step container (libcurvami_h5,
include/curvami_h5.h, over the HDF5 C library); StepWriter.read reads a file
back.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import warnings
from typing import Dict, List, Optional, Sequence

import numpy as np

FIELDS = ("E_XX", "E_YY", "E_XY", "K_XX", "K_YY", "K_XY")
CREASE_COLUMNS = "fold_index,crease_tag,X_ref,Y_ref,s_arc,theta_deg,weight"


def field_codes(fields: Sequence[str]) -> List[int]:
    """Column codes for the optional fields; an unknown name gets a zero column
    and one warning, as in the reference (:82-90)."""
    codes = []
    for f in fields:
        if f in FIELDS:
            codes.append(FIELDS.index(f))
        else:
            warnings.warn(f'unknown output field "{f}" - writing zeros. Valid names: {", ".join(FIELDS)}.')
            codes.append(-1)
    return codes


def planarity(step: np.ndarray):
    """Best-fit plane of the deformed GP cloud (X+u, Y+v, w) by PCA: RMS and
    largest distance from it, and the reference diagonal (:93-129)."""
    if step.shape[0] <= 2:
        return 0.0, 0.0, 0.0
    P = np.stack([step[:, 0] + step[:, 2], step[:, 1] + step[:, 3], step[:, 4]], axis=1)
    mean = P.mean(axis=0)
    D = P - mean
    Cv = D.T @ D / P.shape[0]
    evals, evecs = np.linalg.eigh(Cv)
    n = evecs[:, 0]
    rms = math.sqrt(max(0.0, float(evals[0])))
    pmax = float(np.abs(D @ n).max())
    diag = math.hypot(float(step[:, 0].max() - step[:, 0].min()), float(step[:, 1].max() - step[:, 1].min()))
    return rms, pmax, diag


def crease_table(fold_gps, theta_deg, crease_tags):
    """Rows [fold_index, crease_tag, X_ref, Y_ref, s_arc, theta_deg, weight] per
    crease GP, chained from one of the two mutually farthest GPs to the nearest
    unused one, s_arc at the middle of each GP's weight (:164-201)."""
    rows = []
    for fi, ((gx, gy, w), th) in enumerate(zip(fold_gps, theta_deg)):
        n = len(gx)
        if n == 0:
            continue
        P = np.stack([gx, gy], axis=1)
        d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(axis=2)
        best, start = -1.0, 0
        for i in range(n):                       # first pair (i < j) of largest distance
            j = int(np.argmax(np.where(np.arange(n) > i, d2[i], -1.0)))
            if j > i and d2[i, j] > best:
                best, start = d2[i, j], i
        used = np.zeros(n, bool)
        order = [start]
        used[start] = True
        cur = start
        for _ in range(1, n):
            cand = np.where(used, np.inf, d2[cur])
            nxt = int(np.argmin(cand))
            if not np.isfinite(cand[nxt]):
                break
            used[nxt] = True
            order.append(nxt)
            cur = nxt
        s_acc = 0.0
        for i in order:
            rows.append([float(fi), float(crease_tags[fi]), gx[i], gy[i], s_acc + 0.5 * w[i], th[i], w[i]])
            s_acc += w[i]
    return np.array(rows, dtype=np.float64).reshape(-1, 7)


def step_record(plate, sheet, x, fields: Sequence[str] = (), crease_tags: Optional[Sequence[int]] = None,
                angles=(), is_driver=(), disp=0.0, W_ext=0.0, energies=None, git_commit="unknown"):
    """Datasets and attributes of one exported step, from the device plate."""
    step = plate.export_fields(x, field_codes(fields))
    tags = list(crease_tags) if crease_tags is not None else [-1] * len(sheet.folds)
    th = [plate.fold_angles(f, x) * (180.0 / math.pi) for f in range(len(sheet.folds))]
    creases = crease_table([(f.left.gx, f.left.gy, f.left.w) for f in sheet.folds], th, tags)
    E_mem, E_ben, E_hin = energies if energies is not None else plate.energies(x)
    rms, pmax, diag = planarity(step)
    attrs = {"git_commit": git_commit, "angles": list(angles) or [0.0], "is_driver": list(is_driver) or [0],
             "disp": float(disp), "W_ext": float(W_ext), "E_mem": float(E_mem), "E_ben": float(E_ben),
             "E_hin": float(E_hin), "planar_rms": rms, "planar_max": pmax, "ref_diag": diag,
             "columns": ["X", "Y", "u", "v", "w", "rotX_deg", "rotY_deg"] + list(fields)}
    return step, creases, attrs


_H5 = None
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)
H5_SCALARS = ("disp", "W_ext", "E_mem", "E_ben", "E_hin", "planar_rms", "planar_max", "ref_diag")


def _h5():
    global _H5
    if _H5 is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcurvami_h5.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(path)
        L.cvm_h5_last_error.restype = C.c_char_p
        L.cvm_h5_write_step.argtypes = [C.c_char_p, C.c_int, _d, C.c_int64, C.c_int, _d, C.c_int64, _d, C.c_int,
                                        _i, C.c_int, _d, C.c_char_p]
        L.cvm_h5_read_dataset.argtypes = [C.c_char_p, C.c_char_p, _i64, _d, C.c_int64]
        L.cvm_h5_read_attr.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, _d, C.c_int, _i]
        L.cvm_h5_read_attr_string.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.cvm_h5_list.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _i]
        _H5 = L
    return _H5


def _h5check(st):
    if st != 0:
        raise RuntimeError(_h5().cvm_h5_last_error().decode(errors="replace"))


class StepWriter:
    """Appends steps to an HDF5 file with the reference's dataset names,
    chunking, compression and attributes (Exporter.h:203-322). The file is
    created when it does not exist and opened for writing otherwise, so the
    steps already written survive a later failure."""

    def __init__(self, path: str):
        self.path = path

    def write(self, step: int, data: np.ndarray, creases: np.ndarray, attrs: dict):
        data = np.ascontiguousarray(data, dtype=np.float64)
        creases = np.ascontiguousarray(creases, dtype=np.float64).reshape(-1, 7)
        ang = np.ascontiguousarray(attrs.get("angles", []), dtype=np.float64)
        drv = np.ascontiguousarray(attrs.get("is_driver", []), dtype=np.int32)
        sc = np.array([float(attrs.get(k, 0.0)) for k in H5_SCALARS])
        L = _h5()
        _h5check(L.cvm_h5_write_step(self.path.encode(), int(step), data.ctypes.data_as(_d), data.shape[0],
                                     data.shape[1], creases.ctypes.data_as(_d) if creases.size else None,
                                     creases.shape[0], ang.ctypes.data_as(_d) if ang.size else None, ang.size,
                                     drv.ctypes.data_as(_i) if drv.size else None, drv.size, sc.ctypes.data_as(_d),
                                     str(attrs.get("git_commit", "unknown")).encode()))

    @staticmethod
    def read_dataset(path: str, name: str) -> np.ndarray:
        L = _h5()
        dims = (C.c_int64 * 2)()
        _h5check(L.cvm_h5_read_dataset(path.encode(), name.encode(), dims, None, 0))
        out = np.empty((dims[0], dims[1]))
        _h5check(L.cvm_h5_read_dataset(path.encode(), name.encode(), dims, out.ctypes.data_as(_d), out.size))
        return out

    @staticmethod
    def read_attr(path: str, dataset: str, attr: str) -> np.ndarray:
        L = _h5()
        n = C.c_int(0)
        _h5check(L.cvm_h5_read_attr(path.encode(), dataset.encode(), attr.encode(), None, 0, C.byref(n)))
        out = np.empty(n.value)
        _h5check(L.cvm_h5_read_attr(path.encode(), dataset.encode(), attr.encode(), out.ctypes.data_as(_d),
                                    out.size, C.byref(n)))
        return out

    @staticmethod
    def read_attr_string(path: str, dataset: str, attr: str) -> str:
        buf = C.create_string_buffer(4096)
        _h5check(_h5().cvm_h5_read_attr_string(path.encode(), dataset.encode(), attr.encode(), buf, 4096))
        return buf.value.decode()

    @staticmethod
    def names(path: str) -> List[str]:
        L = _h5()
        need = C.c_int(0)
        _h5check(L.cvm_h5_list(path.encode(), None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _h5check(L.cvm_h5_list(path.encode(), buf, need.value, C.byref(need)))
        return [n for n in buf.value.decode().split(",") if n]

    @classmethod
    def read(cls, path: str):
        """{step: (data, creases or None, attrs)} with the HDF5 attributes."""
        out = {}
        names = cls.names(path)
        for k in names:
            if not (k.startswith("step_") and k.count("_") == 1):
                continue
            s = int(k[5:])
            cr = cls.read_dataset(path, k + "_creases") if (k + "_creases") in names else None
            attrs = {a: float(cls.read_attr(path, k, a)[0]) for a in H5_SCALARS}
            attrs["angles"] = cls.read_attr(path, k, "angles").tolist()
            attrs["is_driver"] = [int(v) for v in cls.read_attr(path, k, "is_driver")]
            attrs["git_commit"] = cls.read_attr_string(path, k, "git_commit")
            out[s] = (cls.read_dataset(path, k), cr, attrs)
        return out
