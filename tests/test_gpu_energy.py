"""Energy conservation on the device without the thermostat (canon off, the
reference's NVE build): the total energy from the device's own observables
kernel (energy.f:25-101 restated) stays within 1e-10 relative of its start
value over long runs at 4 wavefronts per system, for the 512-bead system of
the small configuration and for a batch of uneven systems in one launch.
The runs go through every event kind: cores, bonds, square wells, H-bond
wells with their shoulders, and the resolved outcomes."""
import numpy as np
import pytest

import pdmd

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _drift_ok(eng, sid, e0):
    e = eng.observables(sid).ered
    return abs(e - e0) <= TOL * abs(e0), e


def test_nve_512_beads():
    sysd = pdmd.generate([16] * 8, density=0.01, tstar=0.20, seed=61)
    eng = pdmd.Engine(waves=4)
    try:
        sid = eng.add_system(sysd, seed=3, canon=False)
        e0 = eng.observables(sid).ered
        for _ in range(5):
            eng.advance([sid], [100000])
            ok, e = _drift_ok(eng, sid, e0)
            assert ok, (e, e0, (e - e0) / e0)
        st = eng.stats(sid)
        assert st.status == 0 and st.coll >= 500000
        tally = np.array(st.nevents)
        assert tally[1] > 0 and tally[2] > 0 and tally[3] > 0  # cores, inner and outer bond walls
        assert tally[20:28].sum() > 0  # well events, by resolved outcome
        assert st.nupdates > 0
    finally:
        eng.close()


def test_nve_uneven_batch():
    specs = [([8] * 4, 0.02, 0.15), ([12] * 20, 0.005, 0.5), ([32] * 6, 0.012, 0.3), ([16] * 40, 0.01, 0.2),
             ([24] * 3, 0.018, 0.4), ([8] * 60, 0.008, 0.25)]
    systems = [pdmd.generate(cl, d, t, seed=90 + k) for k, (cl, d, t) in enumerate(specs)]
    eng = pdmd.Engine(waves=4)
    try:
        ids = [eng.add_system(s, seed=10 + k, canon=False) for k, s in enumerate(systems)]
        e0 = [eng.observables(i).ered for i in ids]
        for _ in range(2):
            eng.advance(ids, [100000] * len(ids))
            for sid, ref in zip(ids, e0):
                ok, e = _drift_ok(eng, sid, ref)
                assert ok, (sid, e, ref, (e - ref) / ref)
        assert all(eng.stats(i).status == 0 for i in ids)
    finally:
        eng.close()
