"""The serving loop (cvm_server_*): results equal the direct assembly and are
bit-identical alone and in a concurrent mix; admission respects the HBM budget."""
import threading
import time

import numpy as np
import pytest

import curvami_gpu as G
import serving as S
import workloads as W

pytestmark = pytest.mark.gpu

FLAGS = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS


@pytest.fixture(scope="module")
def fleet():
    sides = [21, 31, 41]
    sheets, plates, _ = S.prepare_sheets(sides)
    reqs = [(sides[i % 3], 0.1 + 0.2 * i, 77 + i) for i in range(12)]
    xs = S.displacements(reqs, sheets)
    return sides, sheets, plates, reqs, xs


def _nnz(plate, flags):
    return plate.nnz_union if flags & G.ELASTIC_FOLDS else plate.nnz_face


def test_alone_equals_direct_and_mix(fleet):
    sides, sheets, plates, reqs, xs = fleet
    srv = G.Server(hbm_budget_bytes=16 << 30, max_inflight=4)
    sid = {n: srv.add_sheet(plates[n]) for n in sides}
    flags = [FLAGS | (G.FP32_TANGENT if i % 4 == 3 else 0) for i in range(len(reqs))]
    # alone: one request at a time
    alone = []
    for i, (n, _, _) in enumerate(reqs):
        K = G.DeviceArray(_nnz(plates[n], flags[i]))
        F = np.empty(3 * plates[n].nn)
        srv.wait(srv.submit(sid[n], xs[i], F, flags[i], K))
        alone.append((F.copy(), K.get()))
        K.free()
        # equals the context's own assembly
        Fd, Kd = plates[n].assemble(xs[i], fp32_tangent=bool(flags[i] & G.FP32_TANGENT))
        assert np.array_equal(F.view(np.int64), Fd.view(np.int64))
        assert np.array_equal(alone[-1][1].view(np.int64), Kd.view(np.int64))
    # mix: everything submitted at once, twice over, four in flight
    Ks, Fs, ids = [], [], []
    for rep in range(2):
        for i, (n, _, _) in enumerate(reqs):
            K = G.DeviceArray(_nnz(plates[n], flags[i]))
            F = np.empty(3 * plates[n].nn)
            ids.append(srv.submit(sid[n], xs[i], F, flags[i], K))
            Ks.append(K); Fs.append(F)
    for r in ids:
        srv.wait(r)
    for j in range(len(ids)):
        i = j % len(reqs)
        assert np.array_equal(Fs[j].view(np.int64), alone[i][0].view(np.int64)), f"request {i}: F differs in the mix"
        assert np.array_equal(Ks[j].get().view(np.int64), alone[i][1].view(np.int64)), f"request {i}: K differs"
    st = srv.stats()
    assert st["completed"] == 3 * len(reqs) and st["inflight"] == 0 and st["pending"] == 0
    # requests admitted together shared tile launches (FP64 and FP32 in separate groups)
    bs = srv.batching_stats()
    print(bs)
    assert bs["shared_tile_launches"] >= 1 and bs["requests_in_shared_launches"] >= 2
    srv.close()


def test_shared_tile_launches_bit_identical(fleet):
    """The same burst with shared tile launches and with every request on its own:
    identical bits, and the shared run did share."""
    sides, sheets, plates, reqs, xs = fleet
    flags = [FLAGS | (G.FP32_TANGENT if i % 3 == 1 else 0) for i in range(len(reqs))]
    out = {}
    for on in (True, False):
        srv = G.Server(hbm_budget_bytes=16 << 30, max_inflight=8)
        srv.set_batching(on)
        sid = {n: srv.add_sheet(plates[n]) for n in sides}
        Fs, Ks, ids = [], [], []
        for i, (n, _, _) in enumerate(reqs):
            K = G.DeviceArray(_nnz(plates[n], flags[i]))
            F = np.empty(3 * plates[n].nn)
            ids.append(srv.submit(sid[n], xs[i], F, flags[i], K))
            Fs.append(F); Ks.append(K)
        for r in ids:
            srv.wait(r)
        out[on] = ([F.copy() for F in Fs], [K.get() for K in Ks], srv.batching_stats())
        for K in Ks:
            K.free()
        srv.close()
    print(out[True][2], out[False][2])
    assert out[True][2]["shared_tile_launches"] >= 1 and out[False][2]["shared_tile_launches"] == 0
    for i in range(len(reqs)):
        assert np.array_equal(out[True][0][i].view(np.int64), out[False][0][i].view(np.int64))
        assert np.array_equal(out[True][1][i].view(np.int64), out[False][1][i].view(np.int64))


def test_budget_admission(fleet):
    sides, sheets, plates, reqs, xs = fleet
    big = max(sides)
    probe = G.Server(hbm_budget_bytes=1 << 40, max_inflight=1)
    s0 = probe.add_sheet(plates[big])
    need = probe.request_bytes(s0, FLAGS, with_K=True)
    probe.close()
    # budget for one request of the largest sheet: never two in flight
    srv = G.Server(hbm_budget_bytes=int(need * 1.5), max_inflight=4)
    sid = {n: srv.add_sheet(plates[n]) for n in sides}
    peak = [0]
    stop = [False]

    def watch():
        while not stop[0]:
            peak[0] = max(peak[0], srv.stats()["inflight"])
            time.sleep(0.0002)

    th = threading.Thread(target=watch)
    th.start()
    outs, ids = [], []
    for i in range(6):
        F = np.empty(3 * plates[big].nn)
        outs.append(F)
        ids.append(srv.submit(sid[big], xs[2], F, FLAGS))
    for r in ids:
        srv.wait(r)
    stop[0] = True
    th.join()
    assert peak[0] <= 1
    for F in outs[1:]:
        assert np.array_equal(F.view(np.int64), outs[0].view(np.int64))
    # one request larger than the whole budget is refused at submit
    tiny = G.Server(hbm_budget_bytes=1 << 20, max_inflight=2)
    t0 = tiny.add_sheet(plates[big])
    with pytest.raises(G.CvmError) as ei:
        tiny.submit(t0, xs[2], np.empty(3 * plates[big].nn), FLAGS)
    assert ei.value.code == G.E_BUDGET
    tiny.close()
    srv.close()


def test_poisson_run_smoke(fleet):
    """A short open-loop run completes every request with ordered percentiles."""
    sides, sheets, plates, reqs, xs = fleet
    srv = G.Server(hbm_budget_bytes=16 << 30, max_inflight=4)
    sid = {n: srv.add_sheet(plates[n]) for n in sides}
    r = S.run_rate(srv, sid, reqs, xs, rate=500.0, seed=9, flags=FLAGS)
    print(r)
    assert r["requests"] == len(reqs) and 0 < r["p50_ms"] <= r["p99_ms"]
    srv.close()


def test_waits_on_a_forgotten_request(fleet):
    """A request is forgotten by the wait that returns it: two threads waiting
    on one request get one result and one refusal (no hang), and a later wait
    on the id is refused."""
    sides, sheets, plates, reqs, xs = fleet
    n = sides[-1]
    srv = G.Server(hbm_budget_bytes=4 << 30, max_inflight=2)
    sid = srv.add_sheet(plates[n])
    F = np.empty(3 * plates[n].nn)
    rid = srv.submit(sid, xs[2], F, FLAGS)
    codes = []

    def waiter():
        try:
            srv.wait(rid)
            codes.append(G.OK)
        except G.CvmError as e:
            codes.append(e.code)

    th = [threading.Thread(target=waiter) for _ in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in th)
    assert sorted(codes) == sorted([G.OK, G.E_INVALID])
    with pytest.raises(G.CvmError) as ei:
        srv.wait(rid)
    assert ei.value.code == G.E_INVALID
    srv.close()
