"""Neighbour lists on the device against the oracle's (nbor.f:23-137
restated, pinned by the reference's own lists in test_oracle_reference.py):
right after set-up every bead's up and down rows hold the oracle's partners
in the same order (the device sorts each row into the order of nbor.f's
cell traversal after its search), the longest row and the totals agree;
after 20000 entries (several list rebuilds at the syncs) the rows agree
again.  Systems of up to 256 beads build their lists with the all-pairs
stencil filter, larger ones through the cell sort; both are covered, at 1
and 4 wavefronts, and random dense systems at 1, 4 and 8."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _oracle_rows(orc, n):
    # the oracle numbers beads from 1, as the reference does
    up = np.array([orc.nb_up(i).size for i in range(1, n + 1)])
    dn = np.array([orc.nb_dn(i).size for i in range(1, n + 1)])
    return up, dn


def _same_rows(eng, sid, orc, n):
    gup, gdn = eng.lists(sid)
    for i in range(1, n + 1):
        assert np.array_equal(gup[i - 1], orc.nb_up(i)), f"up row of bead {i}"
        assert np.array_equal(gdn[i - 1], orc.nb_dn(i)), f"down row of bead {i}"


@pytest.mark.parametrize("chains,density,tstar", [
    ([8] * 4, 0.0101, 0.27),     # 128 beads
    ([12] * 5, 0.02, 0.43),      # 240 beads
    ([4] * 16, 0.03, 0.35),      # 256 beads, the largest of the stencil filter
    ([8] * 9, 0.012, 0.30),      # 288 beads, cell sort
    ([16] * 8, 0.01, 0.20),      # 512 beads, cell sort
    ([8] * 4, 0.3, 0.30),        # 128 beads, dense: a down row longer than 20
    ([8] * 9, 0.25, 0.30),       # 288 beads, dense, cell sort: three rows longer than 20
], ids=["128", "240", "256", "288", "512", "128d", "288d"])
@pytest.mark.parametrize("waves", [1, 4])
def test_list_lengths_match_oracle(param_dir, chains, density, tstar, waves):
    _lists_match(param_dir, chains, density, tstar, waves, 0)


@pytest.mark.parametrize("chains,density,tstar,cap", [
    ([8] * 4, 0.0101, 0.27, 17),  # dense path
    ([8] * 4, 0.3, 0.30, 25),     # dense path, a row longer than 20
    ([8] * 9, 0.25, 0.30, 27),    # cell sort, rows longer than 20
    ([16] * 8, 0.01, 0.20, 18),   # cell sort
], ids=["128-17", "128d-25", "288d-27", "512-18"])
def test_rows_in_order_at_capacities_off_a_multiple_of_4(param_dir, chains, density, tstar, cap):
    """Row capacities that are not a multiple of 4 (the sort then reads rows
    entry by entry instead of 16 bytes at a time)."""
    _lists_match(param_dir, chains, density, tstar, 4, cap)


def _lists_match(param_dir, chains, density, tstar, waves, cap):
    sysd = pdmd.generate(chains, density=density, tstar=tstar, seed=77)
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(sysd, seed=5, list_capacity=cap)
        n = eng.nbeads[sid]
        orc = Oracle(sysd, seed=5, param_dir=param_dir)
        up, dn = _oracle_rows(orc, n)
        st = eng.stats(sid)
        assert st.status == 0
        assert st.max_list == max(up.max(), dn.max())
        assert round(st.avg_up * n) == up.sum() and round(st.avg_dn * n) == dn.sum()
        _same_rows(eng, sid, orc, n)
        eng.advance([sid], [20000])
        orc.run(20000)
        st = eng.stats(sid)
        assert st.status == 0 and st.nupdates > 0
        assert st.nupdates == orc.stats().nupdates
        up, dn = _oracle_rows(orc, n)
        assert round(st.avg_up * n) == up.sum() and round(st.avg_dn * n) == dn.sum()
        _same_rows(eng, sid, orc, n)
    finally:
        eng.close()


@pytest.mark.parametrize("case", range(12))
def test_dense_random_rows_and_events(param_dir, case):
    """Random dense systems (rho* 0.08-0.3, 192-1280 beads, rows of up to
    about 30 entries, many longer than the 20 a lane sorts alone) at 1, 4 and
    8 wavefronts: rows equal to the oracle's in order after set-up and after
    5000 entries, and those 5000 entries equal to the oracle's, times bit
    for bit."""
    rng = np.random.default_rng(123 + 1000 * case)
    nch, L = int(rng.integers(3, 20)), int(rng.choice([8, 12, 16]))
    dens, tstar = float(rng.uniform(0.08, 0.3)), float(rng.uniform(0.15, 0.5))
    canon = bool(case % 2)
    sysd = pdmd.generate([L] * nch, density=dens, tstar=tstar, seed=1000 + case)
    eng = pdmd.Engine(waves=(1, 4, 8)[case % 3])
    try:
        sid = eng.add_system(sysd, seed=7 + case, canon=canon)
        n = eng.nbeads[sid]
        orc = Oracle(sysd, seed=7 + case, param_dir=param_dir, canon=canon)
        _same_rows(eng, sid, orc, n)
        eng.log_enable(sid, 5000)
        eng.advance([sid], [5000])
        g = eng.log_read(sid, 5000)
        ol = orc.run(5000, log_cap=5000)
        assert eng.stats(sid).status == 0
        assert np.array_equal(g["i"], ol["i"]) and np.array_equal(g["j"], ol["j"])
        assert np.array_equal(g["t"].view(np.int64), ol["t"].view(np.int64))
        _same_rows(eng, sid, orc, n)
    finally:
        eng.close()
