"""pdmd_observables_noalloc against pdmd_observables through the C ABI: the
entry point that allocates nothing (its chain tables and results pass
through one fixed device array per GPU) gives the same values bit for bit,
on two calls in a row -- a system of 48 chains, then one of 6 chains, so the
second call finds the first call's longer chain table in the array -- and
again after the systems have advanced."""
import numpy as np
import pytest

import pdmd

pytestmark = pytest.mark.gpu

FIELDS = ("ered", "tred", "sumvel", "ehh_ii", "ehh_ij", "rg_avg", "e2e_avg", "hb_alpha", "hb_ii", "hb_ij")


def _same(a, b):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        if isinstance(x, float):
            assert np.float64(x).view(np.int64) == np.float64(y).view(np.int64), (f, x, y)
        else:
            assert x == y, (f, x, y)


def test_noalloc_observables_equal_allocating_ones():
    big = pdmd.generate([16] * 48, density=0.01, tstar=0.2, seed=31)
    small = pdmd.generate([8] * 6, density=0.01, tstar=0.3, seed=32)
    print("noalloc observables: 48 chains then 6 chains", flush=True)
    eng = pdmd.Engine(waves=4)
    try:
        sb = eng.add_system(big, seed=5)
        ss = eng.add_system(small, seed=6)
        for rnd in range(2):
            if rnd:
                eng.advance([sb, ss], [20000, 20000])
            _same(eng.observables_noalloc(sb), eng.observables(sb))
            _same(eng.observables_noalloc(ss), eng.observables(ss))
        assert eng.observables_noalloc(sb).rg_avg > 0.0 and eng.observables_noalloc(ss).rg_avg > 0.0
    finally:
        eng.close()
