"""Several wavefronts per system: rounds of side-by-side pair events must
commit exactly the serial event sequence.  Each system's event log and final
state are compared bit for bit with the CPU oracle, for 2, 4 and 8
wavefronts, on single systems and on a batch of uneven systems in one launch,
and after a list update forced between events (pdmd_rebuild).
"""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _compare(eng, sid, sysd, seed, n, param_dir, canon=True, orc=None):
    g = eng.log_read(sid, n)
    st = eng.stats(sid)
    assert st.status == 0
    if orc is None:
        orc = Oracle(sysd, seed=seed, param_dir=param_dir, canon=canon)
    o = orc.run(n, log_cap=n)
    assert len(g) == n
    bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                     (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
    assert bad.size == 0, f"event logs diverge at event {bad[0] + 1}: gpu {g[bad[0]]} oracle {o[bad[0]]}"
    gs, os_ = eng.download(sid), orc.state()
    assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64))
    for key in ("identity", "bptnr", "extra", "nptnr", "coltype"):
        assert np.array_equal(gs[key], os_[key]), key
    assert np.array_equal(gs["tim"].view(np.int64), os_["tim"].view(np.int64))


@pytest.mark.parametrize("waves", [2, 4, 8])
@pytest.mark.parametrize("canon", [True, False])
def test_single_system_matches_serial(param_dir, waves, canon):
    n = 12000
    sysd = pdmd.generate([16] * 24, density=0.01, tstar=0.22, seed=31)
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(sysd, seed=77, canon=canon)
        eng.log_enable(sid, n)
        # uneven chunks: rounds end at every event target
        for chunk in (1, 7, 992, 5000, 6000):
            eng.advance([sid], [chunk])
        assert eng.stats(sid).nupdates >= 1, "no list rebuild inside the compared window"
        _compare(eng, sid, sysd, 77, n, param_dir, canon)
    finally:
        eng.close()


@pytest.mark.parametrize("waves", [4])
def test_uneven_batch_matches_serial(param_dir, waves):
    n = 6000
    specs = [([8] * 4, 0.02, 0.15), ([12] * 30, 0.005, 0.5), ([32] * 12, 0.012, 0.3),
             ([16] * 60, 0.01, 0.2), ([24] * 3, 0.018, 0.4), ([8] * 90, 0.008, 0.25)]
    systems = [pdmd.generate(cl, d, t, seed=50 + k) for k, (cl, d, t) in enumerate(specs)]
    eng = pdmd.Engine(waves=waves)
    try:
        ids = [eng.add_system(s, seed=200 + k) for k, s in enumerate(systems)]
        for sid in ids:
            eng.log_enable(sid, n)
        eng.advance(ids, [n] * len(ids))
        for k, (sid, s) in enumerate(zip(ids, systems)):
            _compare(eng, sid, s, 200 + k, n, param_dir)
    finally:
        eng.close()


@pytest.mark.parametrize("waves", [1, 4])
def test_forced_rebuild_matches_oracle(param_dir, waves):
    # pdmd_rebuild between events (mid-interval, after an uneven number of
    # entries) against the oracle's orc_rebuild at the same point
    sysd = pdmd.generate([16] * 20, density=0.01, tstar=0.25, seed=33)
    n1, n2 = 4321, 8000
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(sysd, seed=19)
        eng.advance([sid], [n1])
        up0 = eng.stats(sid).nupdates
        eng.rebuild([sid])
        assert eng.stats(sid).nupdates == up0 + 1 and eng.stats(sid).tfalse == 0.0
        eng.log_enable(sid, n2)
        eng.advance([sid], [n2])
        orc = Oracle(sysd, seed=19, param_dir=param_dir)
        orc.run(n1)
        orc.rebuild()
        _compare(eng, sid, sysd, 19, n2, param_dir, orc=orc)
    finally:
        eng.close()
