"""The C++ host mirror (curvami-official_amd/host/curvami_gpu.hpp), driven like
the reference's main drives Meshless, gives the oracle's pattern and values."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import workloads as W
from _cases import sheet

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "curvami-official_amd", "host", "mirror_check")


@pytest.mark.parametrize("flag_rig", [False, True])
def test_host_mirror_matches_oracle(tmp_path, flag_rig):
    if not os.path.exists(EXE):
        pytest.fail("host/mirror_check not built (make -C curvami-official_amd)")
    sh = W.with_edge_moment_load(sheet("straight21"))
    f = sh.folds[0]
    cr = sh.creases[0]
    x = W.displacement(sh.nn, 1, 1e-3)
    arrays = {
        "meta": [sh.rho, sh.order, f.Kt, f.Ang0, float(flag_rig), float(len(sh.boundaries))],
        "X": sh.X, "Y": sh.Y, "plies": sh.plies.ravel(),
        "crease": np.concatenate([cr.vx, cr.vy, [float(cr.ext_s), float(cr.ext_e)]]),
        "fx": sh.face.gx, "fy": sh.face.gy, "fw": sh.face.w, "x": x,
    }
    for side, g in (("L", f.left), ("R", f.right)):
        for k in ("x", "y", "w", "nx", "ny", "tx", "ty"):
            arrays[side + k] = getattr(g, "g" + k if k in ("x", "y") else k)
    for i, bd in enumerate(sh.boundaries):
        g = bd.gps
        for k in ("x", "y", "w", "nx", "ny", "tx", "ty"):
            arrays[f"b{i}{k}"] = getattr(g, "g" + k if k in ("x", "y") else k)
        arrays[f"b{i}meta"] = [2.0 if bd.kind == "point" else 1.0, float(bd.active), *bd.val, *bd.force, bd.moment]
    for k, v in arrays.items():
        np.asarray(v, dtype=np.float64).tofile(tmp_path / f"{k}.bin")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outer = np.fromfile(tmp_path / "K_outer.bin", np.int32)
    inner = np.fromfile(tmp_path / "K_inner.bin", np.int32)
    val = np.fromfile(tmp_path / "K_values.bin")
    F = np.fromfile(tmp_path / "F.bin")
    orc = O.plate_from_sheet(sh)
    o_outer, o_inner, o_val, o_F = orc.assemble(x, flag_rig=flag_rig)
    assert np.array_equal(outer, o_outer) and np.array_equal(inner, o_inner)
    assert np.abs(val - o_val).max() <= 1e-12 * np.abs(o_val).max()
    assert np.abs(F - o_F).max() <= 1e-10 * np.abs(o_F).max()
    n0 = int(np.diff(orc.neighbours(0, sh.rho)[0])[0])
    assert np.array_equal(np.fromfile(tmp_path / "map0.bin", np.int32), orc.gp_map(0, n0))
    # evaluateConstraints and computeExternalForce through the mirror
    c_o, r_o, col_o, v_o = orc.constraints(x, flag_rig=flag_rig)
    assert np.array_equal(np.fromfile(tmp_path / "J_rows.bin", np.int32), r_o)
    assert np.array_equal(np.fromfile(tmp_path / "J_cols.bin", np.int32), col_o)
    assert np.abs(np.fromfile(tmp_path / "c.bin") - c_o).max() <= 1e-12
    assert np.abs(np.fromfile(tmp_path / "J_vals.bin") - v_o).max() <= 1e-10 * max(1.0, np.abs(v_o).max())
    Fe = orc.external_force()
    assert np.abs(np.fromfile(tmp_path / "F_ext.bin") - Fe).max() <= 1e-11 * np.abs(Fe).max()
