"""Benchmark of the per-Gauss-point hot path on one MI355X.

Workload (one step), on device-resident data: the neighbour search of every
integration domain of the 201 x 201 arc-crease sheet (done again on the
device, as SetupHelper::recomputeShapeFunctions searches before every RKPM
solve), the crease indicators and node sides, the RKPM shape functions (order
2, quartic window, crease visibility), then one residual + block-upper CSC
tangent assembly with the elastic hinge. The assembly plan is the geometry
setup (once per sheet, as initSparsityPattern in the reference) and is
reported separately.

Prints one JSON line. `value` = face Gauss points per second of the step.
"""
import argparse
import ctypes as C
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "curvami-official_amd"))

import curvami_gpu as G  # noqa: E402
import workloads as W  # noqa: E402

_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector / matrix
RHO_FACTOR = 3.75


# ---- algorithmic work per face GP (n = its neighbour count), FP64 ---------------
def sf_flops(n):
    return 1870.0 * n + 5200.0


def sf_bytes(n):
    return 52.0 * n + 16.0


def asm_flops(n):
    return 181.0 * n * (n + 1) / 2.0 + 362.0 * n + 250.0


def asm_bytes(n, nn, ng, nnz, sK=8.0):
    return 44.0 * n + 8.0 + 48.0 * nn / ng + sK * nnz / ng


class Hip:
    """Minimal HIP runtime access for device buffers (plumbing only)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        if self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        return p.value

    def h2d(self, dst, arr):
        if self.L.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def d2h(self, arr, src):
        if self.L.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")

    def d2d(self, dst, src, nbytes):
        if self.L.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3) != 0:
            raise RuntimeError("hipMemcpy D2D failed")

    def free(self, p):
        self.L.hipFree(C.c_void_p(p))

    def sync(self):
        self.L.hipDeviceSynchronize()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(sh, threads, reps=5):
    """The FP64 oracle (restated reference CPU path, OpenMP structure of the
    reference: per-thread F/K copies, guided(8), index-ordered merge) on a
    bounded sample of the same sheet: the face GPs of cell rows 95..104
    (10 of 200 rows, 32,000 GPs, crossed by the arc) plus every fold GP. Timed
    scope as the device step: neighbour search + shape functions of every
    domain, then one assembleSystem; median of `reps` runs after one warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    q = 16                      # 4 x 4 GPs per cell
    ncell = sh.n_side - 1
    lo, hi = 95 * ncell * q, 105 * ncell * q
    sub = W.GPSet(sh.face.gx[lo:hi], sh.face.gy[lo:hi], sh.face.w[lo:hi])
    sample = W.Sheet(sh.name + "_rows95-104", sh.n_side, sh.X, sh.Y, sub, sh.creases, sh.folds, sh.rho,
                     sh.order, sh.kernel, sh.E, sh.nu, sh.t)
    out = {}
    x = W.displacement(sample.nn, 1, 1e-3)
    for T in sorted({1, threads}):
        log(f"cpu baseline: {sub.n} face GPs on {T} thread(s), 1 + {reps} runs")
        p = O.plate_from_sheet(sample, compute=False)
        runs = []
        for r in range(reps + 1):
            t0 = time.perf_counter()
            for d in range(len(p.domains)):
                p.neighbours(d, sample.rho)
            t_nb = time.perf_counter() - t0
            t0 = time.perf_counter()
            for d in range(len(p.domains)):
                st, _, _ = p.shape_functions(d, sample.rho, sample.order, sample.kernel, nthreads=T)
                assert st == 0
            t_sf = time.perf_counter() - t0
            if r == 0:
                assert p.init_assembly() == 0
            t0 = time.perf_counter()
            p.assemble(x, flag_rig=False, nthreads=T)
            t_asm = time.perf_counter() - t0
            if r > 0:
                runs.append((t_nb + t_sf + t_asm, t_nb, t_sf, t_asm))
        runs.sort()
        med = runs[len(runs) // 2]
        out[T] = (sub.n / med[0],) + med[1:]
        p = None
    return out, sub.n


def _step_time(plate, ndom, order, kernel, xd, Fd, Kd, flags, steps, warmup, graph=False):
    """Seconds per step: the neighbour search of every domain again on the
    device and the shape functions of every domain (one call: the shape
    functions formed beside the search's verification, one status read-back),
    then one evaluation (eager, or the captured graph)."""
    def step():
        if plate.recompute_shape_functions(order, kernel):
            raise RuntimeError("neighbour lists changed on unchanged geometry")
        if graph:
            plate.graph_launch()
        else:
            plate.assemble_device(xd.ptr, Fd.ptr, Kd.ptr, flags)
    if graph:
        plate.graph_capture(xd.ptr, Fd.ptr, Kd.ptr, flags)
    for _ in range(warmup):
        step()
    plate.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    plate.synchronize()
    return (time.perf_counter() - t0) / steps


def batched_leg(nsheets, warmup, steps):
    """256 independent sheets (sides U{21..121}, one straight crease each): one
    batched context, one launch per stage, assembly replayed from a captured
    graph, tile shape from the autotuner (tuned on the first 16 sheets). The
    plain FP64 GPU path -- one context per sheet, sheets one after another -- is
    timed on the first 32 sheets for the ratio."""
    import autotune as AT
    log(f"batched: generating {nsheets} sheets")
    sheets = [W.config_batched_sheet(s) for s in range(nsheets)]
    flags64 = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    flags32 = flags64 | G.FP32_TANGENT
    res = {}
    # plain FP64 path, sheet by sheet
    sample = sheets[:min(32, nsheets)]
    tot_gp, tot_t = 0, 0.0
    log(f"batched: plain FP64 path on {len(sample)} sheets, one context each")
    for k, sh in enumerate(sample):
        if k % 8 == 7:
            log(f"batched: plain path, sheet {k + 1}/{len(sample)}")
        p = G.plate_from_sheet(sh)
        x = G.DeviceArray(3 * sh.nn); x.set(W.displacement(sh.nn, 1, 1e-3))
        F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(p.nnz_union)
        tot_t += _step_time(p, len(p.domain_sizes), sh.order, sh.kernel, x, F, K, flags64, steps, 1)
        tot_gp += sh.face.n
        for a in (x, F, K):
            a.free()
        p.close()
    res["plain_fp64_per_sheet_gps"] = tot_gp / tot_t
    res["plain_sample"] = f"first {len(sample)} sheets, {tot_gp} face GPs, one context per sheet, sequential"
    # autotune on a subset
    sub = sheets[:min(16, nsheets)]
    log(f"batched: autotuning the tile shape on {len(sub)} sheets")
    tp = G.batch_from_sheets(sub)
    xs = np.concatenate([W.displacement(sh.nn, 1 + k, 1e-3) for k, sh in enumerate(sub)])
    x = G.DeviceArray(xs.size); x.set(xs)
    F = G.DeviceArray(xs.size); K = G.DeviceArray(tp.nnz_union)
    tune = AT.autotune(tp, x.ptr, F.ptr, K.ptr, flags32, reps=3)
    for a in (x, F, K):
        a.free()
    tp.close()
    tg, um, gr = tune["best"]
    res["autotune"] = {"best": {"tile_gps": tg, "union_max": um, "grid": gr},
                       "table": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}
                                 for r in tune["table"]], "tuned_on": f"first {len(sub)} sheets"}
    # the full batch
    log(f"batched: set-up of the {nsheets}-sheet context")
    t0 = time.perf_counter()
    bp = G.batch_from_sheets(sheets, compute=False)
    bp.set_tile_shape(tg, um, gr)
    for d in range(len(bp.domain_sizes)):
        bp.neighbours(d, 0.0)
    bp.shape_functions_all(sheets[0].order, sheets[0].kernel)
    nf, nu = bp.init_assembly()
    res["setup_s"] = round(time.perf_counter() - t0, 2)
    ng = bp.domain_sizes[bp.face_dom]
    xs = np.concatenate([W.displacement(sh.nn, 1 + k, 1e-3) for k, sh in enumerate(sheets)])
    x = G.DeviceArray(xs.size); x.set(xs)
    F = G.DeviceArray(xs.size); K = G.DeviceArray(nu)
    nd = len(bp.domain_sizes)
    for name, fl in (("mixed", flags32), ("fp64", flags64)):
        log(f"batched: timing the {name} step")
        dt = _step_time(bp, nd, sheets[0].order, sheets[0].kernel, x, F, K, fl, steps, warmup, graph=True)
        # assembly alone (the per-Newton-iteration work), graph replays
        bp.graph_capture(x.ptr, F.ptr, K.ptr, fl)
        bp.graph_launch(); bp.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bp.graph_launch()
        bp.synchronize()
        da = (time.perf_counter() - t0) / steps
        res[name] = {"gps": ng / dt, "ms_per_step": dt * 1e3, "assembly_ms": da * 1e3,
                     "assembly_gps": ng / da}
    res["mixed_over_plain_fp64"] = res["mixed"]["gps"] / res["plain_fp64_per_sheet_gps"]
    res["mixed_over_batched_fp64"] = res["mixed"]["gps"] / res["fp64"]["gps"]
    res["workload"] = (f"{nsheets} sheets, sides U{{21..121}} (mean {np.mean([s.n_side for s in sheets]):.1f}), "
                       f"{ng} face GPs, {bp.npairs[bp.face_dom]} pairs, nnz_K {nu}; step = shape functions "
                       f"(all {nd} domains) + residual + CSC tangent")
    for a in (x, F, K):
        a.free()
    bp.close()
    return res


def large_leg(warmup, steps):
    """1001 x 1001 nodes (16M face GPs), two crossing straight creases and an
    arc; fold 0 a rigid hinge (constraint rows, FP64), folds 1 and 2 elastic.
    Step = shape functions of every domain + residual + tangent (graph replay);
    the constraint rows of the rigid fold are timed separately."""
    log("large: generating and setting up the 1001 x 1001 sheet")
    sh = W.config5(1001)
    t0 = time.perf_counter()
    p = G.plate_from_sheet(sh, compute=False)
    p.set_fold_rigid(0)
    for d in range(len(p.domain_sizes)):
        p.neighbours(d, sh.rho)
    p.shape_functions_all(sh.order, sh.kernel)
    nf, nu = p.init_assembly()
    setup = time.perf_counter() - t0
    ng = sh.face.n
    x = G.DeviceArray(3 * sh.nn); x.set(W.displacement(sh.nn, 1, 1e-3))
    F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(nu)
    nd = len(p.domain_sizes)
    res = {"workload": f"four-vertex sheet 1001x1001 nodes, {ng} face GPs, {p.npairs[p.face_dom]} pairs, "
                       f"nnz_K {nu}, rho 3.75h, creases x=0.5 (rigid), y=0.5, arc r=0.3 (elastic)",
           "setup_s": round(setup, 1)}
    base = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    for name, fl in (("mixed", base | G.FP32_TANGENT), ("fp64", base)):
        log(f"large: timing the {name} step")
        dt = _step_time(p, nd, sh.order, sh.kernel, x, F, K, fl, steps, warmup, graph=True)
        res[name] = {"gps": ng / dt, "ms_per_step": dt * 1e3}
        Fh = F.get()
        if not np.all(np.isfinite(Fh)):
            raise RuntimeError(f"large leg: non-finite residual ({name})")
        res[name]["max_abs_F"] = float(np.abs(Fh).max())
    rows, cols = p.constraint_layout(rigid=False)
    xv = W.displacement(sh.nn, 2, 1e-3)
    p.constraints(xv)
    t0 = time.perf_counter()
    for _ in range(steps):
        p.constraints(xv)
    res["constraint_rows"] = {"rows": p.n_crows, "triplets": p.n_ctrip,
                              "ms_host_in_out": (time.perf_counter() - t0) / steps * 1e3}
    for a in (x, F, K):
        a.free()
    p.close()
    return res


def bspline_leg(n_side, warmup, steps):
    """The main step on the same sheet with the configurations' window, the
    cubic B-spline (the reference has the quartic spline only): neighbour search
    + shape functions + residual + CSC tangent, FP64 and mixed."""
    log(f"bspline: {n_side} x {n_side} arc-crease sheet, cubic B-spline window, set-up")
    sh = W.config2(rho_factor=RHO_FACTOR, n_side=n_side, kernel=1)
    p = G.plate_from_sheet(sh)
    ng = sh.face.n
    x = G.DeviceArray(3 * sh.nn); x.set(W.displacement(sh.nn, 1, 1e-3))
    F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(p.nnz_union)
    res = {"workload": f"arc-crease sheet {n_side}x{n_side} nodes, {ng} face GPs, rho 3.75h, quadratic RKPM, "
                       f"cubic B-spline window, elastic hinge; step as the main one"}
    base = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    for name, fl in (("fp64", base), ("mixed", base | G.FP32_TANGENT)):
        dt = _step_time(p, len(p.domain_sizes), sh.order, sh.kernel, x, F, K, fl, steps, warmup)
        res[name] = {"gps": ng / dt, "ms_per_step": dt * 1e3}
    Fh = F.get()
    assert np.all(np.isfinite(Fh))
    for a in (x, F, K):
        a.free()
    p.close()
    return res


def defaults_leg(n_side, warmup, steps):
    """The same arc-crease sheet at the reference's own defaults: cubic basis
    (order 3) and support 6 h (6 x the 3rd-nearest-node spacing), about 110
    neighbours per GP. Step = neighbour search + shape functions of every
    domain + residual + tangent, as in the main leg, with per-stage device
    times."""
    import dataclasses
    log(f"defaults: {n_side} x {n_side} arc-crease sheet, cubic basis, rho = 6h, set-up")
    sh = dataclasses.replace(W.config2(rho_factor=6.0, n_side=n_side), order=3)
    t0 = time.perf_counter()
    p = G.plate_from_sheet(sh, compute=False)
    for d in range(len(p.domain_sizes)):
        p.neighbours(d, sh.rho)
    p.shape_functions_all(sh.order, sh.kernel)
    _, nu = p.init_assembly()
    setup = time.perf_counter() - t0
    ng = sh.face.n
    x = G.DeviceArray(3 * sh.nn); x.set(W.displacement(sh.nn, 1, 1e-3))
    F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(nu)
    res = {"workload": f"arc-crease sheet {n_side}x{n_side} nodes, {ng} face GPs, {p.npairs[p.face_dom]} pairs, "
                       f"nnz_K {nu}, rho 6h, cubic RKPM, quartic window, elastic hinge",
           "setup_s": round(setup, 1)}
    base = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    for name, fl in (("fp64", base), ("mixed", base | G.FP32_TANGENT)):
        log(f"defaults: timing the {name} step")
        for _ in range(warmup):
            p.recompute_shape_functions(sh.order, sh.kernel)
            p.assemble_device(x.ptr, F.ptr, K.ptr, fl)
        p.synchronize()
        p.set_profiling(True)
        p.reset_profiling()
        t0 = time.perf_counter()
        for _ in range(steps):
            if p.recompute_shape_functions(sh.order, sh.kernel):
                raise RuntimeError("neighbour lists changed on unchanged geometry")
            p.assemble_device(x.ptr, F.ptr, K.ptr, fl)
        p.synchronize()
        dt = (time.perf_counter() - t0) / steps
        st = p.kernel_stats()
        p.set_profiling(False)
        res[name] = {"gps": ng / dt, "ms_per_step": dt * 1e3,
                     "stages_ms": {k: round(v[0] / max(v[1], 1), 3) for k, v in st.items()}}
    for a in (x, F, K):
        a.free()
    p.close()
    return res


def serving_leg(nreq):
    """Poisson request streams over the serving loop: sides U{21..201}, rigid
    fold of a seeded angle plus noise; p50 / p99 latency and throughput per
    rate, FP64 and mixed tangents; the 201 x 201 request alone."""
    import serving as S
    reqs = W.serving_requests(nreq)
    sides = sorted(set(n for n, _, _ in reqs) | {201})
    log(f"serving: preparing {len(sides)} sheets")
    sheets, plates, tprep = S.prepare_sheets(sides)
    xs = S.displacements(reqs, sheets)
    srv = G.Server(hbm_budget_bytes=96 << 30, max_inflight=8)
    sid = {n: srv.add_sheet(plates[n]) for n in sides}
    flags = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
    S.run_rate(srv, sid, reqs[:32], xs[:32], 1e5, 1, flags)                   # warm the slots
    S.run_rate(srv, sid, reqs[:32], xs[:32], 1e5, 1, flags | G.FP32_TANGENT)
    x201 = W.rigid_fold_displacement(sheets[201], 1.0, seed=5, sigma=1e-3)
    out = {"requests_per_rate": nreq, "distinct_sheets": len(sides), "prepare_s": round(tprep, 1),
           "hbm_budget_gb": 96, "slots": 8}
    for name, fl in (("fp64", flags), ("mixed", flags | G.FP32_TANGENT)):
        log(f"serving: {name} tangents, 201 x 201 alone and six request rates")
        p50, p99 = S.size_class_latency(srv, sid, plates, 201, x201, fl)
        rows = [S.run_rate(srv, sid, reqs, xs, r, 100 + r, fl) for r in (50, 100, 200, 500, 1000, 2000)]
        out[name] = {"alone_201_p50_ms": p50, "alone_201_p99_ms": p99,
                     "rates": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}
                               for r in rows]}
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-side", type=int, default=201)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-sheet batched leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 1001 x 1001 leg")
    ap.add_argument("--no-bspline", action="store_true", help="skip the cubic B-spline window leg")
    ap.add_argument("--no-defaults", action="store_true",
                    help="skip the leg at the reference's defaults (cubic basis, rho = 6h)")
    ap.add_argument("--batch-sheets", type=int, default=W.BATCH_SHEETS)
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (the known-size dispatch that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        hip = Hip()
        nb = args.copy_probe << 20
        src, dst = hip.malloc(nb), hip.malloc(nb)
        for _ in range(3):
            hip.d2d(dst, src, nb)
        hip.sync()
        hip.free(src)
        hip.free(dst)
        log(f"copy probe: 3 copies of {args.copy_probe} MiB")
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: sharding a sheet or a batch over "
                                   "several GPUs is designed, not built or measured", "n_gpus": args.gpus}))
        sys.exit(2)

    log(f"main: {args.n_side} x {args.n_side} arc-crease sheet, set-up")
    sh = W.config2(rho_factor=RHO_FACTOR, n_side=args.n_side)
    ng = sh.face.n
    t0 = time.perf_counter()
    p = G.plate_from_sheet(sh, compute=False)
    for d in range(len(p.domain_sizes)):
        p.neighbours(d, sh.rho)
    t_frnn = time.perf_counter() - t0
    ndom = len(p.domain_sizes)
    for d in range(ndom):
        p.shape_functions(d, sh.order, sh.kernel)
    t0 = time.perf_counter()
    nnz_face, nnz_union = p.init_assembly()
    t_plan = time.perf_counter() - t0
    off, ids, _ = p.get_neighbours(0)
    nper = np.diff(off).astype(np.float64)

    hip = Hip()
    x = W.displacement(sh.nn, 1, 1e-3)
    xd = hip.malloc(x.nbytes)
    Fd = hip.malloc(8 * 3 * sh.nn)
    Kd = hip.malloc(8 * nnz_union)
    hip.h2d(xd, x)
    flags = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS

    def step(extra=0, search=True):
        # the search and the shape functions of every domain in one call
        # (cvm_recompute_shape_functions), or the shape functions alone
        if search:
            if p.recompute_shape_functions(sh.order, sh.kernel):
                raise RuntimeError("neighbour lists changed on unchanged geometry")
        else:
            p.shape_functions_all(sh.order, sh.kernel)
        p.assemble_device(xd, Fd, Kd, flags | extra)

    def timed(extra, search=True):
        for _ in range(args.warmup):
            step(extra, search)
        p.synchronize()
        p.set_profiling(True)
        p.reset_profiling()
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(extra, search)
        p.synchronize()
        hip.sync()
        dt_ = (time.perf_counter() - t0) / args.steps
        st_ = p.kernel_stats()
        p.set_profiling(False)
        return dt_, st_

    # plain FP64 path (the reported value), then the mixed path (FP32 tangent);
    # the same two steps without the neighbour search for comparison
    log(f"main: {args.warmup} + {args.steps} FP64 steps")
    dt, stats = timed(0)
    log(f"main: FP64 {dt * 1e3:.3f} ms/step; {args.warmup} + {args.steps} mixed steps")
    dt_mixed, stats_mixed = timed(G.FP32_TANGENT)
    log(f"main: mixed {dt_mixed * 1e3:.3f} ms/step; the same steps without the neighbour search")
    dt_ns, _ = timed(0, search=False)
    dt_ns_mixed, _ = timed(G.FP32_TANGENT, search=False)
    kern_ms = {k: v[0] / max(v[1], 1) for k, v in stats.items()}
    F = np.empty(3 * sh.nn)
    hip.d2h(F, Fd)
    assert np.all(np.isfinite(F))

    # roofline of the dominant kernel (algorithmic work of one launch)
    # shape functions: one launch pair over all domains (the face domain's GPs
    # are all but the 512 fold GPs, whose flops are left out)
    rk_all = stats.get("shape_functions", (0.0, 1))
    tile = stats.get("tile_residual_tangent", (0.0, 1))
    rk_face_ms = rk_all[0] / max(rk_all[1], 1)
    tile_ms = tile[0] / max(tile[1], 1)
    F_sf = float(np.sum(sf_flops(nper)))
    F_asm = float(np.sum(asm_flops(nper)))
    B_asm = float(np.sum(44.0 * nper + 8.0)) + 48.0 * sh.nn + 8.0 * nnz_union
    cands = {
        "shape_functions": (F_sf / (rk_face_ms * 1e-3) / 1e12 if rk_face_ms > 0 else 0.0, rk_face_ms),
        "tile_residual_tangent": (F_asm / (tile_ms * 1e-3) / 1e12 if tile_ms > 0 else 0.0, tile_ms),
    }
    dom_name = max(cands, key=lambda k: cands[k][1])
    achieved = cands[dom_name][0]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        if dom_name in tj and tj[dom_name].get("workload") == sh.name:
            traffic = tj[dom_name]["bytes_per_launch"]
    roofline = {"kernel": dom_name, "bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                "precision": "fp64 VALU (no MFMA; the FP64 vector and matrix peaks are equal on MI355X); "
                             "limited by VALU issue and LDS latency",
                "launch_ms": round(cands[dom_name][1], 4)}
    out = {
        "metric": "Gauss pts/s (shape fns+R+K CSR)",
        "value": ng / dt,
        "unit": "GP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"arc-crease sheet {args.n_side}x{args.n_side} nodes, {ng} face GPs (4x4/cell), "
                               f"rho=3.75h, quadratic RKPM, quartic window, elastic hinge; step = neighbour search "
                               f"+ crease indicators + shape functions (all {ndom} domains) + residual + CSC "
                               f"tangent", "nodes": sh.nn, "face_gps": ng,
                   "pairs": int(p.npairs[0]), "nnz_K": int(nnz_union), "mean_neighbours": float(nper.mean())},
        "stages_ms": {k: round(v, 4) for k, v in kern_ms.items()},
        "mixed": {"value": ng / dt_mixed, "ms_per_step": dt_mixed * 1e3, "speedup_vs_fp64": dt / dt_mixed,
                  "dtype": "f64 shape functions, kinematics, residual, hinge; f32 face tangent",
                  "stages_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in stats_mixed.items()}},
        "without_neighbour_search": {"scope": "shape functions + residual + CSC tangent (neighbour lists and "
                                              "indicators kept from the previous step)",
                                     "fp64": {"value": ng / dt_ns, "ms_per_step": dt_ns * 1e3},
                                     "mixed": {"value": ng / dt_ns_mixed, "ms_per_step": dt_ns_mixed * 1e3}},
        "setup_s": {"first_neighbour_search_host": round(t_frnn, 3), "assembly_plan": round(t_plan, 3)},
        "roofline": roofline,
        "asm_hbm_bytes_algorithmic": B_asm,
    }
    if not args.no_cpu:
        # the CPU share of this process: a GPU box grants 16 host cores per GPU,
        # whatever the machine's core count
        share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = args.cpu_threads or min(16, share)
        res, nsub = cpu_baseline(sh, threads)
        best_T = max(res)
        out["cpu_baseline"] = {"value": res[best_T][0], "unit": "GP/s", "cores": best_T, "kind": "port",
                               "sample": f"oracle (restated reference path, OpenMP) on {nsub} face GPs of cell rows "
                                         f"95-104 of the same sheet + all fold GPs: neighbour search + shape "
                                         f"functions + one assembleSystem, median of 5 runs after one warm-up; "
                                         f"1 thread: {res[1][0]:.0f} GP/s",
                               "seconds_median": {str(T): {"neighbours": round(v[1], 4), "shape_functions":
                                                           round(v[2], 4), "assembly": round(v[3], 4)}
                                                  for T, v in res.items()},
                               "cpu": cpu_model(), "nproc": os.cpu_count(), "cpus_allowed": share,
                               "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    # the reference's default basis (cubic, order 3) on the face domain of the
    # same sheet: shape functions only, device time per launch
    log("main: cubic-basis shape functions (face domain)")
    try:
        p.shape_functions(0, 3, sh.kernel)
        p.synchronize()
        p.set_profiling(True)
        p.reset_profiling()
        ncub = max(3, args.steps // 2)
        for _ in range(ncub):
            p.shape_functions(0, 3, sh.kernel)
        p.synchronize()
        cub = p.kernel_stats().get("rkpm_face", (0.0, 1))
        p.set_profiling(False)
        cub_ms = cub[0] / max(cub[1], 1)
        out["cubic_face_shape_functions"] = {"order": 3, "face_gps": ng, "pairs": int(p.npairs[0]),
                                             "ms": round(cub_ms, 4), "gp_per_s": ng / (cub_ms * 1e-3)}
    except RuntimeError as e:
        out["cubic_face_shape_functions"] = {"error": str(e)}
    for ptr in (xd, Fd, Kd):
        hip.free(ptr)
    p.close()
    if not args.no_bspline:
        out["cubic_bspline_window"] = bspline_leg(args.n_side, args.warmup, args.steps)
    if not args.no_defaults:
        out["reference_defaults"] = defaults_leg(args.n_side, 1, max(3, args.steps // 3))
    if not args.no_batched:
        out["batched"] = batched_leg(args.batch_sheets, args.warmup, max(2, args.steps // 3))
    if not args.no_serving:
        out["serving"] = serving_leg(args.serving_requests)
    if not args.no_large:
        out["large"] = large_leg(args.warmup, max(2, args.steps // 3))
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
