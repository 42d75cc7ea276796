"""Serving loop on the GPU: a request's trajectory is the same alone and in
a mix of other requests (event quanta or device-time slices that cut it at
arbitrary entries, staggered arrivals, admission limited by the memory
budget), bit for bit."""
import numpy as np
import pytest

import pdmd

pytestmark = pytest.mark.gpu


def _sys(chains, seed, tstar=0.2, dens=0.01):
    return pdmd.generate(chains, density=dens, tstar=tstar, seed=seed)


def _alone(system, seed, ncoll):
    e = pdmd.Engine()
    try:
        sid = e.add_system(system, seed=seed)
        e.advance([sid], [ncoll])
        st = e.stats(sid)
        return e.download(sid)["sv"], st.coll, st.status
    finally:
        e.close()


@pytest.mark.parametrize("quantum,slice_ms", [(3000, 0.0), (1000000, 3.0)])
def test_alone_equals_mix(quantum, slice_ms):
    target = _sys([12] * 8, 71)
    others = [(_sys([16] * 6, 72), 9000), (_sys([8] * 20, 73), 15000), (_sys([24] * 10, 74), 4000),
              (_sys([32] * 3, 75, tstar=0.45), 21000), (_sys([4] * 40, 76, dens=0.02), 12000)]
    n = 25000
    sv_alone, coll_alone, status = _alone(target, 5, n)
    assert status == 0 and coll_alone == n

    e = pdmd.Engine()
    srv = pdmd.Server(e, mem_budget_bytes=1 << 36, quantum=quantum, slice_ms=slice_ms)
    try:
        ids = [srv.submit(s, nc, seed=100 + k) for k, (s, nc) in enumerate(others[:3])]
        srv.step()
        srv.step()
        tid = srv.submit(target, n, seed=5)  # arrives while the others are on the device
        srv.step()
        ids += [srv.submit(s, nc, seed=200 + k) for k, (s, nc) in enumerate(others[3:])]
        for _ in range(1000):
            srv.step()
            q, r, _ = srv.pending()
            if q == 0 and r == 0:
                break
        info, sv = srv.result(tid, with_state=True)
        assert info.state == 2 and info.status == 0 and info.done == n
        assert info.t_done >= info.t_admit >= info.t_submit
        assert np.array_equal(sv.view(np.int64), sv_alone.view(np.int64)), "state differs alone vs in the mix"
        for rid in ids:
            inf, _ = srv.result(rid)
            assert inf.state == 2 and inf.done == inf.ncoll
    finally:
        srv.close()
        e.close()


def test_memory_budget_limits_admission():
    systems = [_sys([12] * 6, 80 + k) for k in range(5)]
    d, _ = pdmd.Engine.make_desc(systems[0], 1)
    one = pdmd.lib().pdmd_system_bytes(d)
    e = pdmd.Engine()
    srv = pdmd.Server(e, mem_budget_bytes=int(2.5 * one), quantum=2000)
    try:
        ids = [srv.submit(s, 6000, seed=k) for k, s in enumerate(systems)]
        most = 0
        for _ in range(1000):
            srv.step()
            q, r, _ = srv.pending()
            most = max(most, r)
            if q == 0 and r == 0:
                break
        assert most == 2
        for rid in ids:
            inf, _ = srv.result(rid)
            assert inf.state == 2 and inf.done == 6000
        big = _sys([32] * 200, 90)
        with pytest.raises(pdmd.PdmdError):
            srv.submit(big, 1000, seed=1)
    finally:
        srv.close()
        e.close()


def test_graph_replays_equal_eager():
    # the same chunked advances through replays of one captured launch (with
    # host reads and copies between them) and through direct launches
    systems = [_sys([12] * 6, 81), _sys([8] * 16, 82), _sys([16] * 10, 83, tstar=0.4)]
    chunks = [700, 700, 700, 700, 1300]
    out = []
    for graph in (False, True):
        e = pdmd.Engine(waves=4)
        try:
            if graph:
                e.set_graph_launch(True)
            ids = [e.add_system(s, seed=300 + k) for k, s in enumerate(systems)]
            for sid in ids:
                e.log_enable(sid, sum(chunks))
            for n in chunks:
                e.advance(ids, [n] * len(ids))
                assert all(e.stats(sid).status == 0 for sid in ids)
            out.append([(e.download(sid)["sv"].copy(), e.log_read(sid, sum(chunks))) for sid in ids])
        finally:
            e.close()
    for (sv0, l0), (sv1, l1) in zip(*out):
        assert np.array_equal(sv0.view(np.int64), sv1.view(np.int64))
        assert np.array_equal(l0, l1)
