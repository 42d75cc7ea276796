"""Device error paths, forced from the host: a list capacity too small for
the system's neighbour lists (the rebuild's kErrListCapacity status comes back
as an error of pdmd_system_add), for a 512-bead system (lists built through
the cell sort) and a 128-bead one (lists built by the all-pairs stencil
filter of systems up to 256 beads), a start whose overlapping pair has a
negative collision time (events.f:59-73; kErrNegativeTime, as the oracle
reports it), and calls on ids that do not exist.  The context stays usable
after each error, at 1 and 4 wavefronts."""
import pytest

import pdmd
from test_oracle_invariants import overlapping_start

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chains", [[16] * 8, [8] * 4], ids=["512beads", "128beads"])
@pytest.mark.parametrize("waves", [1, 4])
def test_list_capacity_overflow_is_reported(waves, chains):
    sysd = pdmd.generate(chains, density=0.01, tstar=0.20, seed=11)
    eng = pdmd.Engine(waves=waves)
    try:
        with pytest.raises(RuntimeError, match="status 1"):
            eng.add_system(sysd, seed=3, list_capacity=2)
        sid = eng.add_system(sysd, seed=3)
        eng.advance([sid], [3000])
        st = eng.stats(sid)
        assert st.status == 0 and st.coll == 3000
        with pytest.raises(RuntimeError):
            eng.advance([sid + 7], [10])
        with pytest.raises(RuntimeError):
            eng.advance([sid, sid], [10, 10])
        eng.advance([sid], [1000])
        assert eng.stats(sid).coll == 4000
    finally:
        eng.close()


@pytest.mark.parametrize("waves", [1, 4])
def test_negative_time_at_start_is_reported(waves):
    eng = pdmd.Engine(waves=waves)
    try:
        with pytest.raises(RuntimeError, match="status 2$"):
            eng.add_system(overlapping_start(), seed=1)
        sid = eng.add_system(pdmd.generate([8, 8], density=0.01, tstar=0.2, seed=21), seed=1)
        eng.advance([sid], [2000])
        st = eng.stats(sid)
        assert st.status == 0 and st.coll == 2000
    finally:
        eng.close()
