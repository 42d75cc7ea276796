"""CPU oracle against vectors written by the reference's own subroutines.

The fixtures in tests/golden/ come from oracle/_ref/ref_harness (the
reference's unmodified Fortran compiled by oracle/ref/Makefile, driven by
tests/golden/make_golden.py) on a two-species system (2 chains x 8 residues
and 2 chains x 6 residues, rho* = 0.01, T* = 0.20), built plain, with
-Dno_hbs and with -Dno_h: the scaled parameter
tables (inputinfo.f + scale_down.f), ev_code (make_code.f), the neighbour
lists (nbor_setup.f, cell_link.f, nbor.f), the calendar after events.f,
every pair prediction (core/bond/sqwel/sqshlder/nc_sqwel.f), the energy and
shape observables (energy.f, radgyr.f, end2end.f) and a serial run of
eventdyn.f + partial_events.f + the calendar + the sync pseudo-event in the
reference's order.  Everything is compared bit for bit, observables to
1e-12 relative (their sums run in a different order).
"""
import os

import numpy as np
import pytest

from oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
# (fixture, build flags of the reference harness that wrote it)
CASES = [("ref_2species", {}), ("ref_2species_nohbs", {"no_hbs": True}), ("ref_2species_noh", {"no_h": True})]


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz")))


def _oracle(fx, param_dir, flags):
    sysd = dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=fx["sv0"],
                boxl=float(fx["boxl"]), tstar=float(fx["tstar"]))
    return Oracle(sysd, seed=1, param_dir=param_dir, canon=False, species=fx["species"], **flags)


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
    # 16 wells between chains (the same-chain ones read hp1/hp2 and stay)
    same = fx["chnnum"][:, None] == fx["chnnum"][None, :]
    assert (15 in ref) != bool(flags.get("no_hbs"))
    assert ((ref == 16) & ~same).any() != bool(flags.get("no_h"))
    assert ((ref == 16) & same).any()


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
        orc = _oracle(fx, param_dir, flags)
        n = int(fx["loop_n"])
        keep = fx["loop_i"].size
        log = orc.run(n, log_cap=keep)
        for k in ("i", "j", "type"):
            bad = np.nonzero(log[k] != fx["loop_" + k])[0]
            assert bad.size == 0, f"{name}: {k} differs first at entry {bad[0] + 1}"
        assert np.array_equal(_bits(log["t"]), _bits(fx["loop_t"])), name
        st = orc.state()
        assert np.array_equal(_bits(st["sv"]), _bits(fx["sv1"])), name
        assert np.array_equal(_bits(st["tim"]), _bits(fx["tim1"])), name
        assert np.array_equal(st["nptnr"], fx["nptnr1"]), name
        assert np.array_equal(st["coltype"], fx["coltype1"]), name
        s = orc.stats()
        assert s.nupdates == fx["loop_updates"] and s.nforced == fx["loop_forced"], name
        assert s.tfalse == fx["tfalse1"], name
        _obs_close(orc.observables(), fx["obs_f"][1], fx["obs_i"][1])
        _checks_match(orc.check, fx, name)


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
        assert (got["first_i"], got["first_j"]) == (int(rows[0, 1]), int(rows[0, 2])), (name, sec)
        assert bool(fx["chk_" + sec + "_over"]), name
