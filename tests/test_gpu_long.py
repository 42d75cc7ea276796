"""Long parity windows: twelve seeded random systems, six of up to 256 beads
and six of up to 1536
(thermostat on, so ghost events interleave) advance 3·10^5 calendar entries
each in one batched launch at 2, 4 and 8 wavefronts per system, and every entry of
each event log, times compared bitwise, and the final state equal the CPU
oracle's.  The window spans hundreds of list rebuilds per small system (the
all-pairs stencil filter up to 256 beads, the cell sort above) and many
rounds whose conflicts were replayed rather than cut."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu

N_ENTRIES = 300000


@pytest.mark.parametrize("waves", [2, 4, 8])
def test_long_window_batch(param_dir, waves):
    rng = np.random.default_rng(20261202)
    specs = []
    for k in range(12):
        L = int(rng.choice([4, 8, 12, 16]))
        nch = int(rng.integers(2, (256 if k < 6 else 1536) // (4 * L) + 1))
        specs.append(dict(chains=[L] * nch, density=float(rng.uniform(0.006, 0.02)),
                          tstar=float(rng.uniform(0.15, 0.5)), gen_seed=700 + k, run_seed=900 + k))
    systems = [pdmd.generate(s["chains"], density=s["density"], tstar=s["tstar"], seed=s["gen_seed"]) for s in specs]
    eng = pdmd.Engine(waves=waves)
    try:
        ids = [eng.add_system(sysd, seed=s["run_seed"]) for s, sysd in zip(specs, systems)]
        for sid in ids:
            eng.log_enable(sid, N_ENTRIES)
        eng.advance(ids, [N_ENTRIES] * len(ids))
        for s, sysd, sid in zip(specs, systems, ids):
            st = eng.stats(sid)
            assert st.status == 0, (s, st.status)
            g = eng.log_read(sid, N_ENTRIES)
            orc = Oracle(sysd, seed=s["run_seed"], param_dir=param_dir)
            o = orc.run(N_ENTRIES, log_cap=N_ENTRIES)
            bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                             (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
            assert bad.size == 0, f"{s}: logs diverge at entry {bad[0] + 1}: gpu {g[bad[0]]} oracle {o[bad[0]]}"
            assert st.nupdates == orc.stats().nupdates and st.nupdates > 0
            gs, os_ = eng.download(sid), orc.state()
            assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64)), s
            for key in ("identity", "bptnr", "extra", "nptnr", "coltype"):
                assert np.array_equal(gs[key], os_[key]), (s, key)
            del orc
    finally:
        eng.close()
