"""GPU parity of the HIP event engine against the CPU oracle (same seeded inputs)."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _cfg1(seed=11, tstar=0.20):
    # 8 chains x 16 residues (512 beads), rho* = 0.01, T* = 0.20
    return pdmd.generate([16] * 8, density=0.01, tstar=tstar, seed=seed)


@pytest.fixture(scope="module")
def engine():
    e = pdmd.Engine()
    yield e
    e.close()


def test_predictor_bit_exact(engine, param_dir):
    sysd = _cfg1(seed=3)
    sid = engine.add_system(sysd, seed=5)
    orc = Oracle(sysd, seed=5, param_dir=param_dir)
    ii, jj = [], []
    for i in range(1, orc.N + 1):
        for j in orc.nb_up(i):
            ii.append(i)
            jj.append(int(j))
    rng = np.random.default_rng(0)
    a = rng.integers(1, orc.N + 1, 4000)
    b = rng.integers(1, orc.N + 1, 4000)
    keep = a != b
    ii += list(np.minimum(a, b)[keep])
    jj += list(np.maximum(a, b)[keep])
    tg, yg = engine.predict_pairs(sid, ii, jj)
    to = np.array([orc.predict(int(x), int(y))[0] for x, y in zip(ii, jj)])
    yo = np.array([orc.predict(int(x), int(y))[1] for x, y in zip(ii, jj)])
    assert np.array_equal(yg, yo)
    assert np.array_equal(tg.view(np.int64), to.view(np.int64)), "collision times not bit-identical"
    engine.remove_system(sid)


@pytest.mark.parametrize("canon", [True, False])
def test_event_sequence_first_10k(engine, param_dir, canon):
    sysd = _cfg1(seed=11)
    n = 10000
    sid = engine.add_system(sysd, seed=99, canon=canon)
    engine.log_enable(sid, n)
    engine.advance([sid], [n])
    g = engine.log_read(sid, n)
    orc = Oracle(sysd, seed=99, param_dir=param_dir, canon=canon)
    o = orc.run(n, log_cap=n)
    assert len(g) == n
    first_bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                           (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
    assert first_bad.size == 0, f"event logs diverge at event {first_bad[0] + 1}: gpu {g[first_bad[0]]} oracle {o[first_bad[0]]}"
    gs, os_ = engine.download(sid), orc.state()
    for k in ("identity", "bptnr", "extra", "nptnr", "coltype"):
        assert np.array_equal(gs[k], os_[k]), k
    assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64))
    assert np.array_equal(gs["tim"].view(np.int64), os_["tim"].view(np.int64))
    engine.remove_system(sid)


@pytest.mark.parametrize("waves", [1, 4])
def test_add_system_while_an_advance_runs(param_dir, waves):
    """pdmd_system_add waits for the advance in flight on the context's
    stream before it rewrites the launch's id buffer and descriptor table: a
    system added right after an asynchronous advance of another leaves both
    on the oracle's sequence."""
    a = _cfg1(seed=11)
    b = pdmd.generate([8] * 6, density=0.01, tstar=0.25, seed=12)
    eng = pdmd.Engine(waves=waves)
    try:
        sa = eng.add_system(a, seed=99)
        eng.advance([sa], [6000], wait=False)
        sb = eng.add_system(b, seed=98)
        eng.advance([sa, sb], [2000, 5000])
        ga, gb = eng.download(sa), eng.download(sb)
        assert eng.stats(sa).coll == 8000 and eng.stats(sb).coll == 5000
    finally:
        eng.close()
    oa = Oracle(a, seed=99, param_dir=param_dir)
    oa.run(8000)
    ob = Oracle(b, seed=98, param_dir=param_dir)
    ob.run(5000)
    for g, o in ((ga, oa.state()), (gb, ob.state())):
        assert np.array_equal(g["sv"].view(np.int64), o["sv"].view(np.int64))
        for k in ("identity", "bptnr", "extra", "nptnr", "coltype"):
            assert np.array_equal(g[k], o[k]), k
