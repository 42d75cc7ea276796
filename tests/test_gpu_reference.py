"""The device engine against vectors written by the reference's own
Fortran subroutines (tests/golden, see test_oracle_reference.py): a
two-species system (2 chains x 8 and 2 chains x 6 residues) built plain,
with -Dno_hbs and with -Dno_h, and a dense, cool one where N-C hydrogen bonds
form and break (the worker's resolution of main.F90:1432-1960 with every
shoulder path), no thermostat.  The serial loop as the reference's code ran
it is compared event by event and bit for bit with the device's log (the
first 2e4 entries directly, the whole run through block digests), with the
resolved-coltype tally, the capture-draw count and the final state --
positions, velocities, calendar, identities, bptnr, extra_repuls -- at 1 and
4 wavefronts per system (the two tie fixtures also at 2 and 8); the device's checkover.f / check_nc_int.f kernel
against the reference's own messages on the final and two planted states;
and the device's -Drunr restore (main.F90:240-321) of the final state
against the reference's; the device's neighbour rows after set-up against
the reference's own rows, in nbor.f's order.
"""
import os

import numpy as np
import pytest

import pdmd
from test_oracle_reference import CASES, _checks_match, _obs_close, check_final_state, check_log

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz")))


def _sysd(fx, sv):
    return dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=sv, boxl=float(fx["boxl"]),
                tstar=float(fx["tstar"]))


@pytest.mark.parametrize("waves", [1, 4])
@pytest.mark.parametrize("name,flags", CASES, ids=[c[0] for c in CASES])
def test_device_matches_reference_serial_loop(name, flags, waves):
    _serial_loop(name, flags, waves)


@pytest.mark.parametrize("waves", [2, 8])
@pytest.mark.parametrize("name,flags", [c for c in CASES if c[0] in ("ref_2species_tie", "ref_2species_ptie")],
                         ids=["ref_2species_tie", "ref_2species_ptie"])
def test_device_tie_fixtures_at_other_widths(name, flags, waves):
    """The planted calendar ties and partner ties with rounds of 2 and of
    up to 8 events: every event of a round stamps its adds with its own
    event number, so tied entries added by two events of one round keep the
    reference's last-added-first order."""
    _serial_loop(name, flags, waves)


def _serial_loop(name, flags, waves):
    fx = _fixture(name)
    n = int(fx["loop_n"])
    print(f"reference loop {name}: {n} entries at W={waves}", flush=True)
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(_sysd(fx, fx["sv0"]), seed=int(fx["rng_seed"]), canon=False, species=fx["species"],
                             **flags)
        eng.log_enable(sid, n)
        eng.advance([sid], [n])
        g = eng.log_read(sid, n)
        st = eng.stats(sid)
        gs = eng.download(sid)
        # checkover.f / check_nc_int.f on the device: the final state and the
        # reference's planted states
        _checks_match(lambda sv=None: eng.check(sid, sv), fx, name)
    finally:
        eng.close()
    assert st.status == 0 and st.coll == n
    check_log(g, fx, name)
    check_final_state(gs, fx, name)
    assert list(st.nevents[1:31]) == [int(x) for x in fx["loop_tally"]], name
    assert st.nupdates == fx["loop_updates"] and st.nforced == fx["loop_forced"], name


@pytest.mark.parametrize("name,flags", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("waves", [1, 4])
def test_device_rows_equal_reference_rows(name, flags, waves):
    """Right after set-up, every bead's up and down rows on the device equal
    the rows nbor.f built in the reference's own run, in their order (the
    112-bead fixtures through the all-pairs stencil filter, the 560-bead one
    through the cell sort)."""
    fx = _fixture(name)
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(_sysd(fx, fx["sv0"]), seed=int(fx["rng_seed"]), canon=False, species=fx["species"],
                             **flags)
        gup, gdn = eng.lists(sid)
    finally:
        eng.close()
    up = np.split(fx["nb_flat"], np.cumsum(fx["na_npt"])[:-1])
    dn = np.split(fx["dn_flat"], np.cumsum(fx["nnabdn"])[:-1])
    assert len(gup) == len(up)
    for i in range(len(up)):
        assert np.array_equal(gup[i], up[i]), f"{name}: up row of bead {i + 1}"
        assert np.array_equal(gdn[i], dn[i]), f"{name}: down row of bead {i + 1}"


@pytest.mark.parametrize("name,flags", CASES, ids=[c[0] for c in CASES])
def test_device_observables_match_reference(name, flags):
    """energy.f / radgyr.f / end2end.f on the device against the reference's
    own values before and after the serial loop (1e-12 relative: sums run
    in another order)."""
    fx = _fixture(name)
    eng = pdmd.Engine()
    try:
        sid = eng.add_system(_sysd(fx, fx["sv0"]), seed=int(fx["rng_seed"]), canon=False, species=fx["species"],
                             **flags)
        _obs_close(eng.observables(sid), fx["obs_f"][0], fx["obs_i"][0])
        eng.advance([sid], [int(fx["loop_n"])])
        _obs_close(eng.observables(sid), fx["obs_f"][1], fx["obs_i"][1])
    finally:
        eng.close()


@pytest.mark.parametrize("name,flags", CASES, ids=[c[0] for c in CASES])
def test_device_restart_restore_matches_reference(name, flags):
    """pdmd_system_add with the final state's positions and bptnr restores
    identities, bptnr and shoulder tables as the reference's -Drunr start-up
    does (main.F90:240-321, restated around the reference's repuls_add), and
    check_nc_int.f's counts on the restored state agree."""
    fx = _fixture(name)
    eng = pdmd.Engine()
    try:
        sid = eng.add_system(_sysd(fx, fx["sv1"]), seed=1, canon=False, species=fx["species"], bptnr=fx["bptnr1"],
                             **flags)
        gs = eng.download(sid)
        c = eng.check(sid)
    finally:
        eng.close()
    assert np.array_equal(gs["identity"], fx["rst_identity"]), name
    assert np.array_equal(gs["bptnr"], fx["rst_bptnr"]), name
    assert np.array_equal(gs["extra"].reshape(-1, 4), fx["rst_extra"]), name
    _, boundbad, unboundbad, noss = (int(x) for x in fx["chk_restored_flags"])
    assert (c["boundbad"], c["unboundbad"]) == (boundbad, unboundbad), name
    assert (c["ss_total"] - c["ss_counted"] == 0) == (noss == 0), name
