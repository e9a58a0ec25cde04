"""Tile-shape autotuner for the assembly kernel.

The tile kernel's cost depends on how face Gauss points are grouped: more GPs
per tile share more node pairs (fewer scratch rows, less reduction traffic) but
need more LDS and leave more lanes idle when the neighbour unions are uneven;
the union cap bounds the LDS per tile; a persistent grid trades workgroup
launch overhead against tail imbalance. The tuner times the real evaluation
(residual + tangent, the caller's flags) for each candidate on the caller's
plate and keeps the fastest. Changing the shape never changes the pattern of
K, only the grouping of the sums (round-off level).
"""
from __future__ import annotations

import itertools
import time

import curvami_gpu as G

TILE_GPS = (16, 12, 8)
UNION_MAX = (0, 56)
GRIDS = (0, 2048, 8192)


def _time(plate, x_ptr, F_ptr, K_ptr, flags, reps):
    plate.assemble_device(x_ptr, F_ptr, K_ptr, flags)
    plate.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plate.assemble_device(x_ptr, F_ptr, K_ptr, flags)
    plate.synchronize()
    return (time.perf_counter() - t0) / reps


def autotune(plate, x_ptr, F_ptr, K_ptr, flags, tile_gps=TILE_GPS, union_max=UNION_MAX, grids=GRIDS, reps=3):
    """Returns {"best": (tile_gps, union_max, grid), "ms": ..., "table": [...]}
    and leaves the plate planned with the best shape."""
    table = []
    for tg, um in itertools.product(tile_gps, union_max):
        plate.set_tile_shape(tg, um, 0)
        plate.init_assembly()
        for gr in grids:
            plate.set_tile_shape(tg, um, gr)
            ms = 1e3 * _time(plate, x_ptr, F_ptr, K_ptr, flags, reps)
            table.append({"tile_gps": tg, "union_max": um, "grid": gr, "ms": ms,
                          "ntiles": plate.tile_shape()["ntiles"]})
    best = min(table, key=lambda r: r["ms"])
    plate.set_tile_shape(best["tile_gps"], best["union_max"], best["grid"])
    plate.init_assembly()
    return {"best": (best["tile_gps"], best["union_max"], best["grid"]), "ms": best["ms"], "table": table}
