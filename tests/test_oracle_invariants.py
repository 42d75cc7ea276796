"""Properties of the CPU oracle that the mathematics fixes, for what the
reference-built fixtures leave open (larger systems, the H-bond and shoulder
events, the Andersen thermostat, the project RNG):

* splitmix64 against its published first output, and the oracle's uniform
  deviates against an independent restatement pinned by it;
* the project logarithm against math.log (the ghost times and Maxwell-
  Boltzmann velocities go through it);
* NVE: total energy conserved to 1e-10 relative over 2e5 calendar entries
  with every event type (cores, bonds, wells, H-bonds with shoulders);
* the calendar entry of every bead equals a brute-force minimum over its
  up-list and shoulder partners, after any number of events (to the rounding
  of the free flight since it was predicted);
* event times never run backwards between sync pseudo-events;
* neighbour lists equal a brute-force all-pairs search by the cutoffs;
* Andersen ghosts: momentum of the untouched beads unchanged.
"""
import math

import numpy as np
import pytest

import pdmd
from oracle import Oracle, orc_log, rng_u01

M64 = (1 << 64) - 1


def splitmix64_next(state):
    """One step of SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c):
    returns (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def test_splitmix64_published_value():
    # first output of SplitMix64 from state 0
    assert splitmix64_next(0)[1] == 0xE220A8397B1DCDAF


def test_rng_u01_matches_splitmix64():
    C = 0x5851F42D4C957F2D
    for seed in (0, 1, 12345, 2 ** 63 + 7):
        for stream in (0, 1):
            key = splitmix64_next((seed ^ ((C * (stream + 1)) & M64)) & M64)[1]
            for ctr in (0, 1, 2, 1000, 2 ** 40):
                want = (splitmix64_next((key + ctr) & M64)[1] >> 11) * 2.0 ** -53
                assert rng_u01(seed, stream, ctr) == want


def test_project_log():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000), rng.random(2000) * 1e-300, [1.0, 0.5, 2.0 ** -1074, 1 - 2 ** -53]])
    for x in xs:
        x = float(x)
        if x <= 0:
            continue
        got, want = orc_log(x), math.log(x)
        assert abs(got - want) <= 4 * np.spacing(abs(want)) + 1e-300, (x, got, want)


def _system(chains, seed, tstar=0.2, dens=0.01):
    return pdmd.generate(chains, density=dens, tstar=tstar, seed=seed)


@pytest.mark.parametrize("chains,seed,tstar", [([16] * 8, 11, 0.20), ([8, 24, 12], 12, 0.40), ([4] * 12, 13, 0.15)])
def test_nve_energy_conservation(param_dir, chains, seed, tstar):
    sysd = _system(chains, seed, tstar)
    orc = Oracle(sysd, seed=3, param_dir=param_dir, canon=False)
    e0 = orc.observables().ered
    tally = np.zeros(31, np.int64)
    for _ in range(4):
        orc.run(50000)
        e = orc.observables().ered
        assert abs(e - e0) <= 1e-10 * abs(e0), (e, e0)
    tally = np.array(orc.stats().nevents)
    # the run exercised bonds, cores, wells and resolved H-bond outcomes
    assert tally[1] > 0 and tally[2] > 0 and tally[3] > 0
    assert tally[20:28].sum() > 0


def _brute_tim(orc, i, st):
    best, who = None, None
    N = orc.N
    cands = [int(j) for j in orc.nb_up(i)]
    ex = st["extra"][4 * (i - 1):4 * (i - 1) + 3]
    cands += [int(j) for j in ex if j > i]
    for j in cands:
        t, _ = orc.predict(i, j)
        if t < 1e9 and (best is None or t < best):
            best, who = t, j
    return best, who


@pytest.mark.parametrize("nev", [0, 137, 5000, 40000])
def test_calendar_equals_brute_force(param_dir, nev):
    sysd = _system([16] * 8, 21)
    orc = Oracle(sysd, seed=4, param_dir=param_dir, canon=True)
    orc.run(nev)
    st = orc.state()
    s = orc.stats()
    tf = s.tfalse
    for i in range(1, orc.N + 1):
        best, who = _brute_tim(orc, i, st)
        if best is None or best + tf >= s.interval_max:
            # no event inside the horizon: the "no event" marker
            assert st["tim"][i - 1] >= s.interval_max or st["nptnr"][i - 1] == who
            continue
        # the stored entry was predicted from an earlier time origin: equal
        # up to the rounding of the free flight in between
        assert abs(st["tim"][i - 1] - (best + tf)) <= 1e-10 * s.interval, (i, st["tim"][i - 1], best + tf)
        assert st["nptnr"][i - 1] == who


def test_event_times_monotone_between_syncs(param_dir):
    sysd = _system([12] * 10, 31)
    orc = Oracle(sysd, seed=5, param_dir=param_dir, canon=True)
    n = 100000
    log = orc.run(n, log_cap=n)
    N = orc.N
    t = log["t"]
    seg = np.cumsum(np.r_[0, (log["i"][:-1] == N + 2)])  # a sync starts a new time origin
    for k in np.unique(seg):
        tt = t[seg == k]
        # ltstep (1e-10) of backward slack is what main.F90:547-553 tolerates
        assert np.all(np.diff(tt) >= -1e-10)
    assert np.sum(log["i"] == N + 1) > 0  # ghosts happened
    assert np.sum(log["i"] == N + 2) > 0  # syncs happened


@pytest.mark.parametrize("chains,seed,dens", [([16] * 8, 41, 0.01), ([24] * 6, 42, 0.02), ([8] * 30, 43, 0.005)])
def test_neighbour_lists_brute_force(param_dir, chains, seed, dens):
    sysd = _system(chains, seed, dens=dens)
    orc = Oracle(sysd, seed=1, param_dir=param_dir, canon=False)
    N = orc.N
    st = orc.state()
    x = st["sv"].reshape(N, 6)[:, :3]
    ii, jj = np.triu_indices(N, 1)
    d = x[ii] - x[jj]
    d -= np.round(d)
    r2 = (d * d).sum(1)
    codes = pdmd.pair_codes(sysd, ii + 1, jj + 1)
    rlsq = np.array([0.0] + [orc.table(8, k) for k in range(1, 51)])
    local = ((codes >= 4) & (codes <= 12)) | ((codes >= 17) & (codes <= 26))
    want = local | (r2 <= rlsq[codes])
    listed = np.zeros(N * (N - 1) // 2, bool)
    idx = {}
    for a, b, k in zip(ii, jj, range(len(ii))):
        idx[(a + 1, b + 1)] = k
    for i in range(1, N + 1):
        for j in orc.nb_up(i):
            listed[idx[(i, int(j))]] = True
        for j in orc.nb_dn(i):
            assert i in set(int(q) for q in orc.nb_up(int(j)))
    assert np.array_equal(listed, want), np.argwhere(listed != want)[:5]


def test_ghost_touches_one_bead(param_dir):
    sysd = _system([16] * 8, 51)
    orc = Oracle(sysd, seed=6, param_dir=param_dir, canon=True)
    N = orc.N
    for _ in range(200):
        before = orc.state()["sv"].reshape(N, 6).copy()
        log = orc.run(1, log_cap=1)
        if log["i"][0] != N + 1:
            continue
        after = orc.state()["sv"].reshape(N, 6)
        changed = np.nonzero(np.any(after[:, 3:] != before[:, 3:], axis=1))[0]
        assert list(changed) == [log["j"][0] - 1]
        return
    pytest.skip("no ghost within 200 events")


def test_forced_rebuild_keeps_invariants(param_dir):
    # a list update forced between events (orc_rebuild): the time origin
    # moves to the current time, the lists and the calendar are rebuilt; energy
    # stays conserved, every entry is again the brute-force minimum (now
    # predicted at the new origin, so exactly), and the run goes on
    sysd = _system([16] * 8, 51)
    orc = Oracle(sysd, seed=6, param_dir=param_dir, canon=False)
    e0 = orc.observables().ered
    orc.run(7777)
    s0 = orc.stats()
    assert s0.tfalse > 0.0
    orc.rebuild()
    s1, st = orc.stats(), orc.state()
    assert s1.tfalse == 0.0 and s1.t == s0.t + s0.tfalse and s1.nupdates == s0.nupdates + 1
    for i in range(1, orc.N + 1):
        best, who = _brute_tim(orc, i, st)
        if best is None or best >= s1.interval_max:
            assert st["tim"][i - 1] >= s1.interval_max or st["nptnr"][i - 1] == who
            continue
        assert st["tim"][i - 1] == best and st["nptnr"][i - 1] == who
    assert abs(orc.observables().ered - e0) <= 1e-10 * abs(e0)
    orc.run(30000)
    assert abs(orc.observables().ered - e0) <= 1e-10 * abs(e0)


def overlapping_start(chains=(8, 8), seed=22, depth=0.5):
    """A generated system with chain 1 moved rigidly so that its first
    C-alpha sits inside the hard core of chain 0's first C-alpha (at `depth`
    core diameters) and approaches it; bonds inside each chain stay intact.
    The reference's start-up stops at the pair's negative collision time
    (events.f:59-73)."""
    sysd = pdmd.generate(list(chains), density=0.01, tstar=0.2, seed=seed)
    sv = sysd["sv"].reshape(-1, 6).copy()
    a, b0, b1 = 0, 4 * chains[0], 4 * (chains[0] + chains[1])
    sigma = 3.3 / sysd["boxl"]  # C-alpha core diameter (protein.data), box units
    shift = sv[a, :3] - sv[b0, :3] + np.array([depth * sigma, 0.0, 0.0])
    sv[b0:b1, :3] += shift
    sv[:, :3] -= np.round(sv[:, :3])
    speed = 10.0 * np.abs(sv[:, 3:]).max()
    sv[b0:b1, 3] -= speed  # chain 1 moves towards -x, onto bead a
    out = dict(sysd)
    out["sv"] = sv.ravel()
    return out


def test_overlapping_start_stops_at_negative_time(param_dir):
    # the oracle's start-up reports the negative time (-12), not the
    # negative calendar bin that follows from it; the host half of
    # pdmd_system_add (no device call) accepts the same state, so the device
    # start-up meets the same negative time
    ovl = overlapping_start()
    with pytest.raises(RuntimeError, match=r"\(-12\)"):
        Oracle(ovl, seed=1, param_dir=param_dir)
    d, keep = pdmd.Engine.make_desc(ovl, seed=1, canon=True)
    assert pdmd.lib().pdmd_system_bytes(d) > 0, pdmd.lib().pdmd_last_error()
    Oracle(pdmd.generate([8, 8], density=0.01, tstar=0.2, seed=22), seed=1, param_dir=param_dir)


def test_restore_reports_exhausted_override_slots():
    """A start in which several N-C pairs sit inside their wells around the
    same beads (chains pushed into each other) needs more shoulder overrides
    per bead than the 4 slots the device keeps.  The reference's N x N
    ev_code matrix has no such limit, but its dynamics never produce such a
    state: a non-terminal N-C pair stays inside its well only while the two
    are H-bonded to each other (main.F90:1487-1560 bounces the others), so a
    bead holds at most 3 overrides.  pdmd_system_add reports the overflow
    from its host half, before any device call."""
    d, keep = pdmd.Engine.make_desc(overlapping_start(seed=21), seed=1, canon=True)
    assert pdmd.lib().pdmd_system_bytes(d) < 0
    assert "override slots exhausted" in pdmd.lib().pdmd_last_error().decode()


def nbor_half_stencil():
    """cell_link.f:13-94 (n_wrap = 2): the order of the 62 neighbour-cell
    offsets (dx, dy, dz) that nbor.f scans after a cell's own beads."""
    o, oy = (0, 1, -1, 2, -2), (0, 1, 2)

    def kept(dx, dy, dz):
        return dy != 0 or (dx >= 1 if dz < 1 else dx >= 0)

    order = [(dx, dy, dz) for dz in o[:3] for dy in o[:2] for dx in o[:3] if kept(dx, dy, dz)]
    for iz, dz in enumerate(o):
        for iy, dy in enumerate(oy):
            lo = 3 if (iz < 3 and iy != 2) else 0
            order += [(dx, dy, dz) for dx in o[lo:] if kept(dx, dy, dz)]
    return {d: n + 1 for n, d in enumerate(order)}


def nbor_order_key(cell, pos, M, H, a, b):
    """Where nbor.f finds the pair {a, b}: (cell scanned, the scanned
    cell's own bead's place in its linked list, stencil index (0 for the
    own cell), the other bead's place).  cell[k]: linear cell of
    cell_add.f; pos[k]: 1 + the beads of k's cell with a higher index."""
    nc = M + 4
    ca, cb = cell[a] - 1, cell[b] - 1
    fa = (ca % nc, (ca // nc) % nc, ca // (nc * nc))
    fb = (cb % nc, (cb // nc) % nc, cb // (nc * nc))
    d = tuple(((fb[q] - fa[q]) % M + M // 2) % M - M // 2 for q in range(3))
    if d == (0, 0, 0):
        first, second = (a, b) if pos[a] < pos[b] else (b, a)
        return (cell[a], pos[first], 0, pos[second])
    if d in H:
        return (cell[a], pos[a], H[d], pos[b])
    return (cell[b], pos[b], H[(-d[0], -d[1], -d[2])], pos[a])


@pytest.mark.parametrize("chains,seed,dens", [([16] * 8, 51, 0.01), ([24] * 6, 52, 0.02), ([8] * 30, 53, 0.005)])
def test_row_order_key_reproduces_nbor_order(param_dir, chains, seed, dens):
    """Every neighbour row of the oracle (nbor.f's order, pinned by the
    reference's own lists) comes out of sorting the row's partners by
    nbor_order_key: the tie-break among a row's partners on the device
    follows it."""
    from oracle import lib as orc_lib
    H = nbor_half_stencil()
    assert len(H) == 62 and all(max(abs(x) for x in d) <= 2 for d in H)
    assert all((-d[0], -d[1], -d[2]) not in H for d in H)
    sysd = pdmd.generate(chains, density=dens, tstar=0.25, seed=seed)
    orc = Oracle(sysd, seed=3, param_dir=param_dir, canon=False)
    N, M = orc.N, orc.stats().num_cell - 4
    assert M >= 5
    cell = {k: orc_lib().orc_cell_of(orc.h, k) for k in range(1, N + 1)}
    pos = {}
    members = {}
    for k in range(1, N + 1):
        members.setdefault(cell[k], []).append(k)
    for c, ks in members.items():
        for k in ks:
            pos[k] = 1 + sum(1 for q in ks if q > k)
    n_checked = 0
    for x in range(1, N + 1):
        for row in (orc.nb_up(x), orc.nb_dn(x)):
            want = [int(y) for y in row]
            got = sorted(want, key=lambda y: nbor_order_key(cell, pos, M, H, x, y))
            assert got == want, (x, want, got)
            n_checked += len(want) > 1
    assert n_checked > N // 2
