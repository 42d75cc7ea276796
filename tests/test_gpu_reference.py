"""The device engine against vectors written by the reference's own
Fortran subroutines (tests/golden, see test_oracle_reference.py): a
two-species system (2 chains x 8 and 2 chains x 6 residues) built plain,
with -Dno_hbs and with -Dno_h, no thermostat.  The serial loop of eventdyn.f +
partial_events.f + the calendar + the sync pseudo-event, as the reference's
code ran it, is compared event by event and bit for bit with the device's
log, and the final positions/velocities with the reference's, at 1 and 4
wavefronts per system, and the device's checkover.f / check_nc_int.f
kernel against the reference's own messages on the final and two planted
states.
"""
import os

import numpy as np
import pytest

import pdmd
from test_oracle_reference import _checks_match

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("waves", [1, 4])
@pytest.mark.parametrize("name,flags", [("ref_2species", {}), ("ref_2species_nohbs", {"no_hbs": True}),
                                        ("ref_2species_noh", {"no_h": True})])
def test_device_matches_reference_serial_loop(name, flags, waves):
    fx = dict(np.load(os.path.join(HERE, "golden", name + ".npz")))
    sysd = dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=fx["sv0"],
                boxl=float(fx["boxl"]), tstar=float(fx["tstar"]))
    n = int(fx["loop_n"])
    keep = fx["loop_i"].size
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(sysd, seed=1, canon=False, species=fx["species"], **flags)
        eng.log_enable(sid, keep)
        eng.advance([sid], [n])
        g = eng.log_read(sid, keep)
        st = eng.stats(sid)
        gs = eng.download(sid)
        # checkover.f / check_nc_int.f on the device: the final state and the
        # reference's planted states
        _checks_match(lambda sv=None: eng.check(sid, sv), fx, name)
    finally:
        eng.close()
    assert st.status == 0 and st.coll == n
    for k in ("i", "j", "type"):
        bad = np.nonzero(g[k] != fx["loop_" + k])[0]
        assert bad.size == 0, f"{k} differs first at entry {bad[0] + 1}: device {g[bad[0]]}"
    assert np.array_equal(g["t"].view(np.int64), np.ascontiguousarray(fx["loop_t"]).view(np.int64))
    assert np.array_equal(gs["sv"].view(np.int64), np.ascontiguousarray(fx["sv1"]).view(np.int64))
    assert np.array_equal(gs["nptnr"], fx["nptnr1"]) and np.array_equal(gs["coltype"], fx["coltype1"])
    assert np.array_equal(gs["tim"].view(np.int64), np.ascontiguousarray(fx["tim1"]).view(np.int64))


@pytest.mark.parametrize("name,flags", [("ref_2species", {}), ("ref_2species_nohbs", {"no_hbs": True}),
                                        ("ref_2species_noh", {"no_h": True})])
def test_device_observables_match_reference(name, flags):
    """energy.f / radgyr.f / end2end.f on the device against the reference's
    own values before and after the serial loop (1e-12 relative: sums run
    in another order)."""
    from test_oracle_reference import _obs_close

    fx = dict(np.load(os.path.join(HERE, "golden", name + ".npz")))
    sysd = dict(chain_len=fx["chain_len"], restype=fx["restype"], hp=fx["hp"], sv=fx["sv0"],
                boxl=float(fx["boxl"]), tstar=float(fx["tstar"]))
    eng = pdmd.Engine()
    try:
        sid = eng.add_system(sysd, seed=1, canon=False, species=fx["species"], **flags)
        _obs_close(eng.observables(sid), fx["obs_f"][0], fx["obs_i"][0])
        eng.advance([sid], [int(fx["loop_n"])])
        _obs_close(eng.observables(sid), fx["obs_f"][1], fx["obs_i"][1])
    finally:
        eng.close()
