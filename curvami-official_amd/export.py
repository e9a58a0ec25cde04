"""Step output (Exporter::writeBinaryStep, plate/Exporter.h:27-339).

Per step: the face-GP table [X, Y, u, v, w, rotX_deg, rotY_deg, fields...]
(device: cvm_export_fields), the per-crease dihedral table [fold_index,
crease_tag, X_ref, Y_ref, s_arc, theta_deg, weight] (device angles,
cvm_fold_angles; crease GPs chained nearest-neighbour from one end as in
:164-201), and the step attributes (angles, is_driver, disp, W_ext, E_mem,
E_ben, E_hin, planar_rms, planar_max, ref_diag, git_commit; :93-129, :231-322).

The reference writes these into HDF5 (`/step_NNNNN`, `/step_NNNNN_creases`);
no HDF5 library is available to this build, so StepWriter keeps the same
dataset names, column order and attribute names in a NumPy .npz archive
(`step_NNNNN`, `step_NNNNN_creases`, `step_NNNNN.attrs` as JSON).
"""
from __future__ import annotations

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


class StepWriter:
    """Collects steps under the reference's dataset names and writes them to
    one .npz archive (rewritten in full on every step, so a run that stops
    leaves the steps written so far)."""

    def __init__(self, path: str):
        self.path = path
        self.arrays: Dict[str, np.ndarray] = {}

    def write(self, step: int, data: np.ndarray, creases: np.ndarray, attrs: dict):
        name = f"step_{step:05d}"
        self.arrays[name] = np.ascontiguousarray(data)
        if creases.size:
            self.arrays[name + "_creases"] = np.ascontiguousarray(creases)
            self.arrays[name + "_creases.columns"] = np.array(CREASE_COLUMNS)
        self.arrays[name + ".attrs"] = np.array(json.dumps(attrs))
        tmp = self.path + ".tmp.npz"
        np.savez_compressed(tmp, **self.arrays)
        os.replace(tmp, self.path)

    @staticmethod
    def read(path: str):
        """{step: (data, creases or None, attrs)}."""
        out = {}
        with np.load(path, allow_pickle=False) as z:
            for k in z.files:
                if k.startswith("step_") and k.count("_") == 1 and not k.endswith(".attrs"):
                    s = int(k[5:])
                    cr = z[k + "_creases"] if (k + "_creases") in z.files else None
                    out[s] = (z[k], cr, json.loads(str(z[k + ".attrs"])))
        return out
