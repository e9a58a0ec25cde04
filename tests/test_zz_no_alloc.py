"""cvm_assemble_noalloc (the allocation-free entry point, static device scratch)
against cvm_assemble with device pointers: bit-identical residual and tangent,
FP64 and FP32 tangents, on two calls in a row, the second on a smaller sheet
(the scratch keeps the first call's contents); a sheet whose workspace exceeds
the scratch is refused with CVM_E_BUDGET before any launch."""
import numpy as np
import pytest

import curvami_gpu as G
import workloads as W
from _cases import sheet

pytestmark = pytest.mark.gpu

FLAGS = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS


def _eval(p, sh, xv, flags, noalloc):
    x = G.DeviceArray(3 * sh.nn); F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(p.nnz_union)
    x.set(xv)
    if noalloc:
        p.assemble_noalloc(x.ptr, F.ptr, K.ptr, flags)
    else:
        p.assemble_device(x.ptr, F.ptr, K.ptr, flags)
    p.synchronize()
    out = F.get(), K.get()
    for a in (x, F, K):
        a.free()
    return out


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_matches_assemble(fp32):
    flags = FLAGS | (G.FP32_TANGENT if fp32 else 0)
    big, small = sheet("arc41"), sheet("straight21")
    pb, ps = G.plate_from_sheet(big), G.plate_from_sheet(small)
    xb = W.displacement(big.nn, 61, 1e-3)
    xs = W.displacement(small.nn, 62, 1e-3)
    ref_b = _eval(pb, big, xb, flags, noalloc=False)
    ref_s = _eval(ps, small, xs, flags, noalloc=False)
    for rep in range(2):
        got_b = _eval(pb, big, xb, flags, noalloc=True)       # first call: the larger sheet
        got_s = _eval(ps, small, xs, flags, noalloc=True)     # second call: a smaller one
        for got, ref, nm in ((got_b, ref_b, "arc41"), (got_s, ref_s, "straight21")):
            assert np.array_equal(got[0].view(np.int64), ref[0].view(np.int64)), f"{nm} rep {rep}: F"
            assert np.array_equal(got[1].view(np.int64), ref[1].view(np.int64)), f"{nm} rep {rep}: K"
    # residual only after a full call leaves nothing stale in the result
    x = G.DeviceArray(3 * small.nn); F = G.DeviceArray(3 * small.nn)
    x.set(xs)
    ps.assemble_noalloc(x.ptr, F.ptr, 0, G.RESIDUAL | G.ELASTIC_FOLDS)
    ps.synchronize()
    assert np.array_equal(F.get().view(np.int64), ref_s[0].view(np.int64))
    x.free(); F.free()
    pb.close(); ps.close()


def test_noalloc_refuses_large_sheet():
    sh = W.serving_sheet(121)
    p = G.plate_from_sheet(sh)
    x = G.DeviceArray(3 * sh.nn); F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(p.nnz_union)
    with pytest.raises(G.CvmError) as e:
        p.assemble_noalloc(x.ptr, F.ptr, K.ptr, FLAGS)
    assert e.value.code == G.E_BUDGET and "static scratch holds" in str(e.value)
    print(str(e.value))
    for a in (x, F, K):
        a.free()
    p.close()


def test_noalloc_with_natural_curvature():
    """The natural-curvature field formed at cvm_init_assembly is read, not
    formed, by the allocation-free evaluation: same bits as cvm_assemble."""
    import dataclasses
    import math
    base = sheet("straight21")
    sh = dataclasses.replace(base, natural_curvature=(0.37, 0.12, math.radians(30.0), -1.0))
    p = G.plate_from_sheet(sh)
    xv = W.displacement(sh.nn, 63, 1e-3)
    ref = _eval(p, sh, xv, FLAGS, noalloc=False)
    got = _eval(p, sh, xv, FLAGS, noalloc=True)
    assert np.array_equal(got[0].view(np.int64), ref[0].view(np.int64))
    assert np.array_equal(got[1].view(np.int64), ref[1].view(np.int64))
    plain = G.plate_from_sheet(base)
    assert not np.array_equal(_eval(plain, base, xv, FLAGS, noalloc=True)[0], got[0])
    p.close(); plain.close()
