"""Step output on the device: the exporter's face-GP columns
(cvm_export_fields) against the oracle's restatement, and a progressive run
of the reference's smoke example writing every step through export.StepWriter.

Bars: X, Y, u, v, w within 1e-12 absolute; strain and curvature columns
within 1e-10 of each column's largest value (curvatures carry the second
derivatives' ~1e-11 round-off next to the crease); rotations within 1e-9
degrees; each step's stored energies equal to the run's records and the
crease table's angles equal to the device angles.
"""
import math
import os

import numpy as np
import pytest

import continuation as CS
import curvami_gpu as G
import export as EX
import oracle as O
import progressive as PR
import workloads as W
from _cases import RHO_FACTOR
from _progressive_oracle import smoke_case

pytestmark = pytest.mark.gpu
INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")


@pytest.mark.parametrize("kind", ["random", "rigid_fold"])
def test_export_columns_match_oracle(kind):
    sh = W.config1(rho_factor=RHO_FACTOR)
    orc = O.plate_from_sheet(sh, nthreads=8)
    gpu = G.plate_from_sheet(sh)
    x = W.displacement(sh.nn, 9, 2e-3) if kind == "random" else W.rigid_fold_displacement(sh, 0.7, seed=1, sigma=1e-4)
    codes = [0, 1, 2, 3, 4, 5, -1]
    ref = orc.export_columns(sh.face.gx, sh.face.gy, x, sh.rho, codes)
    got = gpu.export_fields(x, codes)
    assert got.shape == ref.shape
    e = np.abs(got - ref)
    print(f"{kind}: max |d| X..w {e[:, :5].max():.2e}, rot {e[:, 5:7].max():.2e} deg, fields {e[:, 7:].max():.2e}")
    scale = np.maximum(1.0, np.abs(ref[:, 7:]).max(axis=0))
    assert e[:, :5].max() <= 1e-12 and np.all(e[:, 7:].max(axis=0) <= 1e-10 * scale) and e[:, 5:7].max() <= 1e-9
    assert np.all(got[:, 13] == 0.0)


def test_progressive_smoke_writes_steps(tmp_path):
    cfg, ms = smoke_case(INPUTS)
    sh = ms.sheet
    plate = G.plate_from_sheet(sh)
    solver = CS.ContinuationSolver()
    solver.tol_force = float(cfg["solver"]["tol_force"])
    solver.tol_energy = float(cfg["solver"]["tol_energy"])
    be = PR.DeviceBackend(plate, sh, solver, exact_hinge=True,
                          theta_star=[fi.sign * fi.angle_target for fi in ms.fold_info])
    lines, static = PR.lines_from_config(ms, cfg["creasing"])
    out = str(tmp_path / "smoke_steps.h5")
    writer = EX.StepWriter(out)
    tags = [fi.crease_tag for fi in ms.fold_info]

    def on_step(rec, x):
        data, creases, attrs = EX.step_record(plate, sh, x, cfg["outputs"], tags,
                                              energies=(rec.E_mem, rec.E_ben, rec.E_hin))
        writer.write(rec.step, data, creases, attrs)

    x = np.zeros(3 * sh.nn)
    recs = PR.run(be, x, lines, static, PR.CreasingConfig.from_json(cfg["creasing"]), on_step=on_step)
    steps = EX.StepWriter.read(out)
    assert sorted(steps) == [r.step for r in recs]
    data, creases, attrs = steps[recs[-1].step]
    assert data.shape == (sh.face.n, 7 + len(cfg["outputs"]))
    assert attrs["E_hin"] == recs[-1].E_hin
    th = plate.fold_angles(0, x) * 180.0 / math.pi
    assert creases.shape == (sh.folds[0].left.n, 7) and np.allclose(np.sort(creases[:, 5]), np.sort(th), rtol=0,
                                                                      atol=1e-12)
    # the last step's columns are the device export of the final state
    assert np.array_equal(data, plate.export_fields(x, EX.field_codes(cfg["outputs"])))
    print(f"wrote {len(steps)} steps; planar_max {attrs['planar_max']:.3e}, mean |theta| "
          f"{np.mean(np.abs(creases[:, 5])):.4f} deg")
