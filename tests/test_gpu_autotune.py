"""The per-system autotuner on the device: pdmd_autotune_systems probes
every system at 1, 2, 4 and the most wavefronts that fit, each timed by the
system's own device clock, then the list rebuild's coarse cell at factors 2,
3 and 4, and keeps each system's fastest settings.  The probes are ordinary
committed entries: the whole event sequence (probes included) and the final
state stay bit for bit the oracle's, which knows nothing of the tuning."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def test_autotune_systems_keeps_the_sequence(param_dir):
    systems = [pdmd.generate([8] * 4, 0.0101, 0.27, seed=41), pdmd.generate([16] * 8, 0.012, 0.30, seed=42),
               pdmd.generate([17] * 20, 0.01, 0.25, seed=43)]
    probe, more = 3000, 8000
    total = 7 * probe + more
    eng = pdmd.Engine()
    try:
        ids = [eng.add_system(s, seed=70 + k) for k, s in enumerate(systems)]
        for sid in ids:
            eng.log_enable(sid, total)
        w, cf, rates = eng.autotune_systems(ids, probe)
        print("autotune:", list(w), list(cf), np.round(rates / 1e3, 1).tolist(), flush=True)
        eng.advance(ids, [more] * len(ids))
        out = []
        for k, sid in enumerate(ids):
            st = eng.stats(sid)
            assert st.status == 0 and st.coll == total
            assert st.waves == w[k] and st.cell_factor == cf[k]
            assert st.device_ms > 0.0 and st.rebuilds > 0 and 0.0 < st.rebuild_ms < st.device_ms
            out.append((eng.log_read(sid, total), eng.download(sid)))
    finally:
        eng.close()
    for k in range(len(systems)):
        assert w[k] in (1, 2, 4, 5, 6, 7, 8) and cf[k] in (2, 3, 4)
        assert rates[k, 0] > 0.0 and rates[k, 2] > 0.0
    for k, s in enumerate(systems):
        orc = Oracle(s, seed=70 + k, param_dir=param_dir)
        o = orc.run(total, log_cap=total)
        g, gs = out[k]
        bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
        assert bad.size == 0, f"system {k}: logs differ at entry {bad[0] + 1}"
        assert np.array_equal(gs["sv"].view(np.int64), orc.state()["sv"].view(np.int64))
