"""Restart from a saved state, the reference's -Drunr start (main.F90:240-321):
the H-bond partners read from the .bptnr record turn the bonded N/C
identities into 5/8, and every N-C pair inside its well gets its shoulder
overrides back (repuls_add.f).  The saved state is written the way
config.f writes it (positions moved to the current time, velocities,
bptnr) after a CPU run long enough for H-bonds to form; the device restores
a system from it and so does the CPU oracle, and the runs that follow are
compared event by event and bit for bit, with the restored identities and
shoulder tables."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _saved_state(param_dir):
    sysd = pdmd.generate([16] * 24, density=0.02, tstar=0.15, seed=91)
    orc = Oracle(sysd, seed=5, param_dir=param_dir, canon=True)
    orc.run(600000)
    st, tf = orc.state(), orc.stats().tfalse
    sv = st["sv"].reshape(-1, 6).copy()
    sv[:, :3] += sv[:, 3:] * tf  # config.f:22-39 writes the positions at the current time
    saved = dict(sysd, sv=sv.ravel())
    return saved, st["bptnr"].copy()


def test_restart_restore_matches_oracle(param_dir):
    saved, bptnr = _saved_state(param_dir)
    assert (bptnr > 0).sum() >= 4, "the saved state holds too few H-bonds to exercise the restore"
    n = 20000
    orc = Oracle(saved, seed=11, param_dir=param_dir, canon=True, bptnr=bptnr)
    o0 = orc.state()
    eng = pdmd.Engine()
    try:
        sid = eng.add_system(saved, seed=11, canon=True, bptnr=bptnr)
        g0 = eng.download(sid)
        # the restore itself: identities 5/8 on the bonded pairs, shoulders
        for key in ("identity", "bptnr", "extra"):
            assert np.array_equal(g0[key], o0[key]), key
        assert np.isin(g0["identity"], [5, 8]).sum() == (bptnr > 0).sum()
        assert (g0["extra"] > 0).any()
        chk = eng.check(sid)
        assert chk["overlaps"] == 0 and chk["bonds_long"] == 0 and chk["bonds_short"] == 0
        eng.log_enable(sid, n)
        eng.advance([sid], [n])
        g = eng.log_read(sid, n)
        assert eng.stats(sid).status == 0
        o = orc.run(n, log_cap=n)
        bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                         (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
        assert bad.size == 0, f"event logs diverge at event {bad[0] + 1}: gpu {g[bad[0]]} oracle {o[bad[0]]}"
        gs, os_ = eng.download(sid), orc.state()
        assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64))
        for key in ("identity", "bptnr", "extra", "nptnr", "coltype"):
            assert np.array_equal(gs[key], os_[key]), key
    finally:
        eng.close()
