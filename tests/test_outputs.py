"""Step output and the restart record on the CPU.

- The .res record: the byte layout of RestartManager::save
  (core/include/curvami/RestartManager.h:14-40) written by hand with struct
  and compared byte for byte; a round trip; every refusal of load (:45-116).
- The exporter's face-GP columns (plate/Exporter.h:47-91) restated on the
  oracle and pinned by closed forms: a rigid rotation about X by alpha gives
  rotX = alpha (degrees), rotY = 0, the rotated positions and zero strains and
  curvatures; a stretch (1 + e) along X gives E_XX = e + e^2 / 2.
- The crease table's chaining and arc length (:164-201) on a shuffled
  straight crease, and the planarity attributes (:93-129) of a flat and of a
  folded point cloud.
"""
import math
import os
import struct

import numpy as np
import pytest

import curvami_gpu as G
import export as EX
import oracle as O
import workloads as W
from _cases import RHO_FACTOR


def test_restart_bytes_match_the_reference_layout(tmp_path):
    x = np.array([1.5, -2.0, 3.25]); d = np.array([0.0, 1e-300, -7.0])
    p = str(tmp_path / "a.res")
    G.restart_save(p, 12, 0.375, 0.1, -0.02, 4.5, x, d)
    expect = (b"CURVRES\0" + struct.pack("<i", 1) + struct.pack("<i", 12) + struct.pack("<4d", 0.375, 0.1, -0.02, 4.5)
              + struct.pack("<i", 3) + x.astype("<f8").tobytes() + d.astype("<f8").tobytes())
    assert open(p, "rb").read() == expect
    r = G.restart_load(p, expected_ns=3)
    assert (r["k"], r["lam"], r["ds"], r["dl_old"], r["W_ext"]) == (12, 0.375, 0.1, -0.02, 4.5)
    assert np.array_equal(r["x"], x) and np.array_equal(r["dU_old"], d)


def test_restart_refusals(tmp_path):
    x = np.arange(6.0)
    good = str(tmp_path / "g.res")
    G.restart_save(good, 3, 1.0, 0.5, 0.25, 0.0, x, x)
    raw = open(good, "rb").read()

    def bad(name, data, match, expected_ns=-1):
        p = str(tmp_path / name)
        open(p, "wb").write(data)
        with pytest.raises(G.CvmError, match=match):
            G.restart_load(p, expected_ns)

    bad("magic.res", b"NOTARES\0" + raw[8:], "not a Curvami restart record")
    bad("version.res", raw[:8] + struct.pack("<i", 2) + raw[12:], "format version 2")
    bad("header.res", raw[:20], "truncated in its header")
    bad("arrays.res", raw[:-8], "truncated in its state arrays")
    bad("ns.res", raw[:48] + struct.pack("<i", 0) + raw[52:], "non-positive DOF count")
    nan = bytearray(raw); nan[52:60] = struct.pack("<d", float("nan"))
    bad("nan.res", bytes(nan), "non-finite")
    with pytest.raises(G.CvmError, match="holds 6 DOFs but this problem has 9"):
        G.restart_load(good, expected_ns=9)
    with pytest.raises(G.CvmError, match="differ in size"):
        G.restart_save(str(tmp_path / "z.res"), 0, 0, 0, 0, 0, x, x[:3])


@pytest.fixture(scope="module")
def c1():
    sh = W.config1(rho_factor=RHO_FACTOR)
    return sh, O.plate_from_sheet(sh, nthreads=4)


def test_export_columns_rigid_rotation_and_stretch(c1):
    sh, orc = c1
    a = 0.35
    x = np.zeros(3 * sh.nn)             # rotation by a about the X axis: (X, Y, 0) -> (X, Y cos a, Y sin a)
    x[1::3] = sh.Y * math.cos(a) - sh.Y
    x[2::3] = sh.Y * math.sin(a)
    cols = orc.export_columns(sh.face.gx, sh.face.gy, x, sh.rho, codes=list(range(6)) + [-1])
    assert np.array_equal(cols[:, 0], sh.face.gx) and np.array_equal(cols[:, 1], sh.face.gy)
    assert np.abs(cols[:, 2]).max() < 1e-12
    assert np.abs(cols[:, 3] - (sh.face.gy * math.cos(a) - sh.face.gy)).max() < 1e-12
    assert np.abs(cols[:, 4] - sh.face.gy * math.sin(a)).max() < 1e-12
    assert np.abs(cols[:, 5] - math.degrees(a)).max() < 1e-9 and np.abs(cols[:, 6]).max() < 1e-9
    assert np.abs(cols[:, 7:13]).max() < 1e-9 and np.all(cols[:, 13] == 0.0)
    e = 0.01
    x = np.zeros(3 * sh.nn)
    x[0::3] = e * sh.X
    cols = orc.export_columns(sh.face.gx, sh.face.gy, x, sh.rho, codes=[0, 1, 2])
    assert np.abs(cols[:, 7] - (e + 0.5 * e * e)).max() < 1e-11
    assert np.abs(cols[:, 8:10]).max() < 1e-11


def test_crease_table_and_planarity():
    rng = np.random.default_rng(4)
    n = 9
    s = (np.arange(n) + 0.5) / n
    perm = rng.permutation(n)
    gx, gy, w = 0.2 + 0.6 * s[perm], 0.5 + 0.0 * s[perm], np.full(n, 0.6 / n)
    th = np.arange(n, dtype=float)[perm]
    tab = EX.crease_table([(gx, gy, w)], [th], [7])
    assert tab.shape == (n, 7) and np.all(tab[:, 0] == 0) and np.all(tab[:, 1] == 7)
    assert np.allclose(np.abs(np.diff(tab[:, 2])), 0.6 / n)            # walked end to end
    assert np.allclose(tab[:, 4], (np.arange(n) + 0.5) * 0.6 / n)      # s at weight midpoints
    flat = np.zeros((50, 7)); flat[:, 0] = rng.random(50); flat[:, 1] = rng.random(50)
    rms, pmax, diag = EX.planarity(flat)
    assert rms < 1e-15 and pmax < 1e-15 and diag == pytest.approx(math.hypot(np.ptp(flat[:, 0]), np.ptp(flat[:, 1])))
    folded = flat.copy(); folded[:, 4] = np.where(flat[:, 0] > 0.5, 0.1, 0.0)
    rms, pmax, _ = EX.planarity(folded)
    assert rms > 1e-3 and pmax >= rms


def test_hdf5_step_container_layout(tmp_path):
    """The step container (libcurvami_h5) writes the reference's layout
    (Exporter.h:203-322), checked with h5dump, an independent reader: datasets
    /step_NNNNN and /step_NNNNN_creases, chunked (min(rows, 1000), cols) and
    deflate-6 compressed, the crease "columns" string, the step attributes with
    their types; a second step appends to the file; writing an existing step
    fails; values read back bit for bit."""
    import shutil
    import subprocess
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    path = str(tmp_path / "steps.h5")
    rng = np.random.default_rng(0)
    data = rng.standard_normal((2500, 10))
    creases = rng.standard_normal((40, 7))
    attrs = {"git_commit": "abc123", "angles": [0.5, -1.25], "is_driver": [1, 0], "disp": 0.1, "W_ext": 2.0,
             "E_mem": 1e-8, "E_ben": 2e-6, "E_hin": 3e-5, "planar_rms": 0.01, "planar_max": 0.02, "ref_diag": 1.4}
    w = EX.StepWriter(path)
    w.write(3, data, creases, attrs)
    w.write(4, data[:10], np.zeros((0, 7)), dict(attrs, angles=[], is_driver=[]))
    with pytest.raises(RuntimeError):
        w.write(3, data, creases, attrs)
    head = subprocess.run([h5dump, "-H", "-p", path], capture_output=True, text=True, check=True).stdout
    for s in ('DATASET "step_00003"', 'DATASET "step_00003_creases"', 'DATASET "step_00004"',
              "DATASPACE  SIMPLE { ( 2500, 10 ) / ( 2500, 10 ) }", "CHUNKED ( 1000, 10 )", "COMPRESSION DEFLATE { LEVEL 6 }",
              'ATTRIBUTE "angles"', "H5T_ARRAY { [2] H5T_IEEE_F64LE }", 'ATTRIBUTE "is_driver"',
              "H5T_ARRAY { [2] H5T_STD_I32LE }", 'ATTRIBUTE "git_commit"', 'ATTRIBUTE "columns"', "CHUNKED ( 40, 7 )"):
        assert s in head, s
    for a in ("disp", "W_ext", "E_mem", "E_ben", "E_hin", "planar_rms", "planar_max", "ref_diag"):
        assert f'ATTRIBUTE "{a}"' in head
    assert 'DATASET "step_00004_creases"' not in head
    vals = subprocess.run([h5dump, "-a", "/step_00003_creases/columns", path], capture_output=True, text=True,
                          check=True).stdout
    assert "fold_index,crease_tag,X_ref,Y_ref,s_arc,theta_deg,weight" in vals
    got = EX.StepWriter.read(path)
    assert sorted(got) == [3, 4]
    d, c, a = got[3]
    assert np.array_equal(d, data) and np.array_equal(c, creases)
    assert a["angles"] == [0.5, -1.25] and a["is_driver"] == [1, 0] and a["git_commit"] == "abc123"
    assert all(a[k] == attrs[k] for k in ("disp", "W_ext", "E_mem", "E_ben", "E_hin", "planar_rms"))
    assert got[4][1] is None and got[4][2]["angles"] == [0.0] and got[4][2]["is_driver"] == [0]
    # a step without Gauss points: an empty table, stored contiguous
    w.write(5, np.zeros((0, 10)), np.zeros((0, 7)), attrs)
    d5, c5, a5 = EX.StepWriter.read(path)[5]
    assert d5.shape == (0, 10) and c5 is None and a5["disp"] == attrs["disp"]
