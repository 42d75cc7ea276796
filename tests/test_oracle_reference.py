"""CPU oracle against vectors written by the reference's own subroutines.

The fixtures in tests/golden/ come from oracle/_ref/ref_harness (the
reference's unmodified Fortran compiled by oracle/ref/Makefile, driven by
tests/golden/make_golden.py) on a two-species system (2 chains x 8 residues
and 2 chains x 6 residues): rho* = 0.01, T* = 0.20, built plain, with
-Dno_hbs and with -Dno_h, rho* = 0.30, T* = 0.18 plain, where N-C hydrogen
bonds form and break, and a planted-tie start (each species' first chain and
a copy translated by half a box, positions on a 2^-20 grid in a 64 A box) in
which copy and original share collision times bit for bit, so that the
reference's LIFO bucket lists (add_tbin.f:28-32, main.F90:521-545) order the
tied entries.  They hold the scaled parameter tables
(inputinfo.f + scale_down.f), ev_code (make_code.f), the neighbour lists
(nbor_setup.f, cell_link.f, nbor.f), the calendar after events.f, every pair
prediction (core/bond/sqwel/sqshlder/nc_sqwel.f), the energy and shape
observables (energy.f, radgyr.f, end2end.f), a serial run of the event loop
in the reference's order -- every pair event through the worker's resolution
of main.F90:1432-1960 around the reference's repuls_check, repuls_check_3,
check_sigma, eventdyn, bumpoff, repuls_add, repuls_del_a/b and
partial_events, the calendar and the sync and output pseudo-events -- with
the resolved-coltype tally, the final state, check_nc_int.f and checkover.f
on it, and the -Drunr restore of that state (main.F90:240-321) with
check_nc_int.f on the result.  Everything is compared bit for bit,
observables to 1e-12 relative (their sums run in a different order).
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
# (fixture, build flags of the reference harness that wrote it)
CASES = [("ref_2species", {}), ("ref_2species_nohbs", {"no_hbs": True}), ("ref_2species_noh", {"no_h": True}),
         ("ref_2species_hb", {}), ("ref_2species_tie", {}), ("ref_2species_ch2", {"chaptype2": True}),
         ("ref_2species_ptie", {"chaptype2": True}), ("ref_560", {})]
# coltypes of the worker's N-C resolution (main.F90:1487-1937, eventdyn.f)
HB_TYPES = (14, 15, 16, 23, 24, 25, 26)
LOG_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("t", "<f8")])
# residues of the first species block of every fixture (2 chains x 8)
L1_CHAINS = 16


def block_digests(log, block=1000):
    """8-byte BLAKE2b digests of consecutive blocks of log entries packed as
    tests/golden/make_golden.py packs the reference's."""
    rec = np.zeros(len(log), LOG_DTYPE)
    for k in ("i", "j", "type", "t"):
        rec[k] = log[k]
    raw = rec.tobytes()
    step = block * LOG_DTYPE.itemsize
    return np.array([np.frombuffer(hashlib.blake2b(raw[k:k + step], digest_size=8).digest(), "<u8")[0]
                     for k in range(0, len(raw), step)], np.uint64)


def check_log(log, fx, name):
    """Event by event over the kept window, block digests over the rest."""
    keep = fx["loop_i"].size
    for k in ("i", "j", "type"):
        bad = np.nonzero(log[k][:keep] != fx["loop_" + k])[0]
        assert bad.size == 0, f"{name}: {k} differs first at entry {bad[0] + 1}: {log[bad[0]]}"
    assert np.array_equal(_bits(log["t"][:keep]), _bits(fx["loop_t"])), name
    dg = block_digests(log)
    bad = np.nonzero(dg != fx["loop_digest"])[0]
    assert dg.size == fx["loop_digest"].size and bad.size == 0, \
        f"{name}: log differs in block {bad[0] if bad.size else '?'} of {fx['loop_digest'].size}"


def check_final_state(st, fx, name):
    assert np.array_equal(_bits(st["sv"]), _bits(fx["sv1"])), name
    assert np.array_equal(_bits(st["tim"]), _bits(fx["tim1"])), name
    assert np.array_equal(st["nptnr"], fx["nptnr1"]), name
    assert np.array_equal(st["coltype"], fx["coltype1"]), name
    assert np.array_equal(st["identity"], fx["identity1"]), name
    assert np.array_equal(st["bptnr"], fx["bptnr1"]), name
    assert np.array_equal(np.asarray(st["extra"]).reshape(-1, 4), fx["extra1"]), name


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz")))


def _oracle(fx, param_dir, flags, seed=1):
    sysd = dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=fx["sv0"],
                boxl=float(fx["boxl"]), tstar=float(fx["tstar"]))
    return Oracle(sysd, seed=seed, param_dir=param_dir, canon=False, species=fx["species"], **flags)


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, param_dir):
    name, flags = request.param
    fx = _load(name)
    return fx, _oracle(fx, param_dir, flags), flags


def test_tables(case):
    fx, orc, _ = case
    for w, key in enumerate(("sigma_sq", "welldia_sq", "ep_sqrt", "shlddia_sq")):
        got = np.array([[orc.table(w, k, kk) for kk in range(1, 29)] for k in range(1, 29)])
        assert np.array_equal(_bits(got), _bits(fx[key])), key
    got = np.array([[orc.table(6, r, c) for c in range(1, 51)] for r in range(1, 4)])
    assert np.array_equal(_bits(got), _bits(fx["ev_param"]))
    got = np.array([[orc.table(7, r, c) for c in range(9, 29)] for r in range(1, 6)])
    assert np.array_equal(_bits(got), _bits(fx["sqz610"][:, 8:]))
    got = np.array([orc.table(8, k) for k in range(1, 51)])
    assert np.array_equal(_bits(got), _bits(fx["rlsq"]))
    st = orc.stats()
    assert st.sig_max_all == fx["sig_max_all"] and st.hdelr == fx["hdelr"]
    assert st.interval == fx["interval"] and st.interval_max == fx["interval_max"]
    assert st.sortsize == fx["sortsize"] and st.num_cell == int(fx["num_cell"])
    ident = orc.state()["identity"]
    assert np.array_equal(ident, fx["identity"])
    bm = np.array([orc.table(5, int(t)) for t in ident])
    assert np.array_equal(_bits(bm), _bits(fx["bm"]))


def test_ev_code(case):
    fx, orc, flags = case
    N = orc.N
    got = np.array([[orc.ev_code(a, b) for b in range(1, N + 1)] for a in range(1, N + 1)], np.int32)
    ref = fx["ev_code"]
    bad = np.argwhere(got != ref)
    assert bad.size == 0, f"{len(bad)} codes differ, first (a,b)={bad[0] + 1}: oracle {got[tuple(bad[0])]} ref {ref[tuple(bad[0])]}"
    # what each build flag removes: -Dno_hbs every code 15, -Dno_h the code
    # 16 wells between chains (the same-chain ones read hp1/hp2 and stay),
    # -Dchaptype=2 every code but 1 inside a second-species chain and the
    # 16 wells between two second-species chains
    same = fx["chnnum"][:, None] == fx["chnnum"][None, :]
    assert (15 in ref) != bool(flags.get("no_hbs"))
    assert ((ref == 16) & ~same).any() != bool(flags.get("no_h"))
    assert ((ref == 16) & same).any()
    sp2 = np.arange(N) >= 4 * L1_CHAINS
    inner2 = same & sp2[:, None] & sp2[None, :]
    assert (ref[inner2] != 1).any() != bool(flags.get("chaptype2"))
    cross2 = ~same & sp2[:, None] & sp2[None, :]
    if flags.get("chaptype2"):
        assert not ((ref == 16) & cross2).any()


def test_neighbour_lists_in_reference_order(case):
    fx, orc, _ = case
    N = orc.N
    up = np.split(fx["nb_flat"], np.cumsum(fx["na_npt"])[:-1])
    dn = np.split(fx["dn_flat"], np.cumsum(fx["nnabdn"])[:-1])
    for i in range(1, N + 1):
        assert np.array_equal(orc.nb_up(i), up[i - 1]), f"up list of bead {i}"
        assert np.array_equal(orc.nb_dn(i), dn[i - 1]), f"down list of bead {i}"


def test_calendar_after_events(case):
    fx, orc, _ = case
    st = orc.state()
    assert np.array_equal(_bits(st["tim"]), _bits(fx["tim0"]))
    assert np.array_equal(st["nptnr"], fx["nptnr0"])
    assert np.array_equal(st["coltype"], fx["coltype0"])


def test_every_pair_prediction(case):
    fx, orc, _ = case
    N = orc.N
    k = 0
    for i in range(1, N):
        for j in range(i + 1, N + 1):
            t, ty = orc.predict(i, j)
            rt = fx["pred_tij"][k]
            assert np.float64(t).view(np.int64) == np.float64(rt).view(np.int64), (i, j, t, rt)
            if rt < 1e9:
                assert ty == fx["pred_type"][k], (i, j)
            k += 1
    assert k == fx["pred_tij"].size


def _obs_close(o, f, iv):
    got = np.array([o.ered, o.tred, o.sumvel, o.ehh_ii, o.ehh_ij, o.rg_avg, o.e2e_avg])
    np.testing.assert_allclose(got, f, rtol=1e-12, atol=1e-300)
    assert (o.hb_alpha, o.hb_ii, o.hb_ij) == tuple(int(x) for x in iv)


def test_observables_initial(case):
    fx, orc, _ = case
    _obs_close(orc.observables(), fx["obs_f"][0], fx["obs_i"][0])


def test_serial_loop(param_dir):
    # fresh oracles: the module fixture's ones have been queried only
    for name, flags in CASES:
        fx = _load(name)
        orc = _oracle(fx, param_dir, flags, seed=int(fx["rng_seed"]))
        n = int(fx["loop_n"])
        log = orc.run(n, log_cap=n)
        check_log(log, fx, name)
        check_final_state(orc.state(), fx, name)
        s = orc.stats()
        assert s.nupdates == fx["loop_updates"] and s.nforced == fx["loop_forced"], name
        assert s.tfalse == fx["tfalse1"], name
        assert list(s.nevents[1:31]) == [int(x) for x in fx["loop_tally"]], name
        assert s.rng_ctr[1] == fx["loop_draws"], name
        _obs_close(orc.observables(), fx["obs_f"][1], fx["obs_i"][1])
        _checks_match(orc.check, fx, name)


def test_tie_fixture_holds_ties_in_lifo_order():
    """The planted-tie fixture's loop commits entries of equal time back to
    back, the later-added one first: after events.f's adds in index order
    (events.f:111-122) the copy's bead (index + 32) precedes the original."""
    fx = _load("ref_2species_tie")
    N = fx["identity"].size
    tim = fx["tim0"][:N]
    assert (tim[:32] == tim[32:64]).all()
    t, i = fx["loop_t"], fx["loop_i"]
    eq = np.nonzero((t[1:] == t[:-1]) & (i[1:] <= N) & (i[:-1] <= N))[0]
    assert eq.size >= 8
    assert sum(int(i[k]) == int(i[k + 1]) + 32 for k in eq) >= 4


def test_hbond_fixture_exercises_the_resolution():
    """The H-bond fixture's loop runs every branch of the worker's N-C
    resolution: captures refused (14), shoulder exits/entries (15, 24, 25,
    26), wells left with shoulders removed (16), end-bead capture draws (23
    and the draw count), and bonds stand at the end."""
    fx = _load("ref_2species_hb")
    tally = fx["loop_tally"]
    for ty in HB_TYPES:
        assert tally[ty - 1] > 0, f"no coltype {ty} in the H-bond fixture"
    assert sum(int(tally[ty - 1]) for ty in (20, 21) + HB_TYPES) >= 300
    assert int(fx["loop_draws"]) > 0
    assert (fx["bptnr1"] > 0).sum() >= 2
    assert (fx["identity1"] >= 5).sum() >= 2 and (fx["identity1"] <= 8).sum() > 0


def test_restart_restore(param_dir):
    """-Drunr restore (main.F90:240-321) of each fixture's final state:
    identities, bptnr, extra_repuls and ev_code as the harness's restatement
    around the reference's repuls_add left them, and check_nc_int.f's counts
    on the restored state."""
    for name, flags in CASES:
        fx = _load(name)
        sysd = dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=fx["sv1"],
                    boxl=float(fx["boxl"]), tstar=float(fx["tstar"]))
        orc = Oracle(sysd, seed=1, param_dir=param_dir, canon=False, species=fx["species"], bptnr=fx["bptnr1"],
                     **flags)
        st = orc.state()
        assert np.array_equal(st["identity"], fx["rst_identity"]), name
        assert np.array_equal(st["bptnr"], fx["rst_bptnr"]), name
        assert np.array_equal(np.asarray(st["extra"]).reshape(-1, 4), fx["rst_extra"]), name
        N = orc.N
        got = np.array([[orc.ev_code(a, b) for b in range(1, N + 1)] for a in range(1, N + 1)], np.int32)
        assert np.array_equal(got, fx["rst_ev_code"]), name
        _, boundbad, unboundbad, noss = (int(x) for x in fx["chk_restored_flags"])
        c = orc.check()
        assert (c["boundbad"], c["unboundbad"]) == (boundbad, unboundbad), name
        assert (c["ss_total"] - c["ss_counted"] == 0) == (noss == 0), name
    fx = _load("ref_2species_hb")
    assert (fx["rst_extra"][:, 3] > 0).sum() >= 2 and (fx["rst_ev_code"] >= 40).sum() >= 8


def planted_states(fx):
    """The two states the reference's checkover.f ran on (ref_harness.F90):
    bead 5 put 1e-3 box lengths off bead N-3 on each axis, bead 10 moved by
    0.05 box lengths along x."""
    sv = np.array(fx["sv1"], np.float64).reshape(-1, 6)
    N = sv.shape[0]
    p1 = sv.copy()
    p1[4, :3] = sv[N - 4, :3] + 1.0e-3
    p2 = sv.copy()
    p2[9, 0] = sv[9, 0] + 5.0e-2
    return {"planted1": p1.ravel(), "planted2": p2.ravel()}


def _checks_match(check, fx, name):
    """Counts by kind and first pair of a check function against the
    reference's checkover.f / check_nc_int.f messages."""
    got = check()
    over, boundbad, unboundbad, noss = (int(x) for x in fx["chk_final_flags"])
    assert (got["overlaps"] + got["bonds_long"] + got["bonds_short"] > 0) == bool(over), name
    assert fx["chk_final"].shape[0] == got["overlaps"] + got["bonds_long"] + got["bonds_short"], name
    assert got["boundbad"] == boundbad and got["unboundbad"] == unboundbad, name
    assert got["ss_total"] - got["ss_counted"] == 0 and noss == 0, name
    for sec, sv in planted_states(fx).items():
        got = check(sv)
        rows = fx["chk_" + sec]
        want = [int((rows[:, 0] == k).sum()) for k in range(3)]
        assert [got["overlaps"], got["bonds_long"], got["bonds_short"]] == want, (name, sec, got, want)
        first = (int(rows[0, 1]), int(rows[0, 2])) if rows.shape[0] else (0, 0)
        assert (got["first_i"], got["first_j"]) == first, (name, sec)
        assert bool(fx["chk_" + sec + "_over"]) == (rows.shape[0] > 0), name
