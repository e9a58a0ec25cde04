"""Natural curvature (Kinematics::NaturalCurvature, plate/Kinematics.h:9-45,
:114-115) on the device against the oracle: a mountain band across the
straight-crease sheet and a valley band across the arc-crease sheet.

Bars as for the plain plate: F_int to 1e-10 relative (inf-norm), K to 1e-12 of
max|K|, energies to 1e-12 relative, the FP32 tangent to 1e-5 of max|K|. The
band set after the assembly plan (cvm_set_natural_curvature forms the field at
once) gives the same bits as the band set before it, switching it off gives the
plain plate's bits back, and a graph captured without it is refused.
"""
import dataclasses
import math

import numpy as np
import pytest

import curvami_gpu as G
import oracle as O
import workloads as W
from _cases import sheet

pytestmark = pytest.mark.gpu

BANDS = {"straight21": (0.37, 0.12, math.radians(30.0), -1.0),
         "arc41": (0.61, 0.08, math.radians(60.0), 1.0)}


@pytest.mark.parametrize("name", sorted(BANDS))
def test_residual_tangent_energies_match_oracle(name):
    base = sheet(name)
    sh = dataclasses.replace(base, natural_curvature=BANDS[name])
    orc = O.plate_from_sheet(sh)
    gpu = G.plate_from_sheet(sh)
    x = W.displacement(sh.nn, 1, 1e-3)
    _, _, Kr, Fr = orc.assemble(x, flag_rig=False)
    Fg, Kg = gpu.assemble(x, elastic_folds=True)
    ef = np.abs(Fg - Fr).max() / np.abs(Fr).max()
    ek = np.abs(Kg - Kr).max() / np.abs(Kr).max()
    Er, Eg = np.array(orc.energies(x)), np.array(gpu.energies(x))
    ee = np.abs(Eg - Er).max() / np.abs(Er).max()
    _, K32 = gpu.assemble(x, elastic_folds=True, fp32_tangent=True)
    e32 = np.abs(K32 - Kr).max() / np.abs(Kr).max()
    print(f"{name} band {BANDS[name]}: F rel {ef:.3e}, K rel {ek:.3e}, energies rel {ee:.3e}, FP32 K {e32:.3e}")
    assert ef <= 1e-10 and ek <= 1e-12 and ee <= 1e-12 and e32 <= 1e-5
    # the band changes the answer (bending share of F_int)
    plain = O.plate_from_sheet(base)
    assert np.abs(Fr - plain.internal_force(x, flag_rig=False)).max() > 1e-5 * np.abs(Fr).max()


def test_set_after_plan_and_off_again():
    name = "straight21"
    base = sheet(name)
    gpu = G.plate_from_sheet(base)                   # plan built without the band
    x = W.displacement(base.nn, 2, 1e-3)
    F0, K0 = gpu.assemble(x)
    xd, Fd = G.DeviceArray(x.size), G.DeviceArray(3 * base.nn)
    Kd = G.DeviceArray(K0.size)
    xd.set(x)
    flags = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    gpu.graph_capture(xd.ptr, Fd.ptr, Kd.ptr, flags)
    gpu.set_natural_curvature(True, *BANDS[name])
    F1, K1 = gpu.assemble(x)
    with pytest.raises(G.CvmError) as ei:
        gpu.graph_launch()
    assert ei.value.code == G.E_STATE
    ref = G.plate_from_sheet(dataclasses.replace(base, natural_curvature=BANDS[name]))
    F2, K2 = ref.assemble(x)
    assert np.array_equal(F1.view(np.int64), F2.view(np.int64))
    assert np.array_equal(K1.view(np.int64), K2.view(np.int64))
    gpu.set_natural_curvature(False)
    F3, K3 = gpu.assemble(x)
    assert np.array_equal(F3.view(np.int64), F0.view(np.int64))
    assert np.array_equal(K3.view(np.int64), K0.view(np.int64))
    with pytest.raises(G.CvmError) as ei:
        gpu.set_natural_curvature(True, 0.5, 0.0, 0.1, -1.0)
    assert ei.value.code == G.E_INVALID
    for a in (xd, Fd, Kd):
        a.free()


def test_shared_serving_launch_with_bands():
    """Sheets with different bands (and one without) evaluated in one shared tile
    launch of the serving loop: each result is the context's own assembly, bit for bit."""
    specs = [("straight21", BANDS["straight21"]), ("arc41", BANDS["arc41"]), ("straight21", None)]
    plates, xs = [], []
    for k, (name, band) in enumerate(specs):
        sh = dataclasses.replace(sheet(name), natural_curvature=band)
        plates.append(G.plate_from_sheet(sh))
        xs.append(W.displacement(sh.nn, 40 + k, 1e-3))
    flags = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    srv = G.Server(hbm_budget_bytes=8 << 30, max_inflight=4)
    sid = [srv.add_sheet(p) for p in plates]
    outs, Ks, ids, which = [], [], [], []
    for rep in range(4):                      # a burst: requests queue up and are admitted together
        for k, p in enumerate(plates):
            F = np.empty(3 * p.nn)
            K = G.DeviceArray(p.nnz_union)
            ids.append(srv.submit(sid[k], xs[k], F, flags, K))
            outs.append(F); Ks.append(K); which.append(k)
    for r in ids:
        srv.wait(r)
    bs = srv.batching_stats()
    print(bs)
    assert bs["requests_in_shared_launches"] >= 2
    ref = [p.assemble(xs[k]) for k, p in enumerate(plates)]
    for j, k in enumerate(which):
        assert np.array_equal(outs[j].view(np.int64), ref[k][0].view(np.int64))
        assert np.array_equal(Ks[j].get().view(np.int64), ref[k][1].view(np.int64))
        Ks[j].free()
    srv.close()
