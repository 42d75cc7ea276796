"""Randomised parity sweep: many systems drawn at random (chain counts and
lengths, one or two species, density, temperature, thermostat on or off,
the -Dno_hbs / -Dno_h build flags) advance together in one batched launch,
at 1, 2, 4 and 8 wavefronts per system, and each one's event log (2·10^4
entries) and final state must equal the CPU oracle's bit for bit.  The draws are seeded, so a
failure names a reproducible system."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _draw(rng, k, max_chains):
    L = int(rng.choice([4, 6, 8, 12, 16, 24, 32]))
    nch = int(rng.integers(2, max_chains + 1))
    return dict(chains=[L] * nch, density=float(rng.uniform(0.005, 0.03)), tstar=float(rng.uniform(0.10, 0.60)),
                canon=bool(rng.random() < 0.7), no_hbs=bool(rng.random() < 0.2), no_h=bool(rng.random() < 0.2),
                gen_seed=1000 + k, run_seed=5000 + k, species=None)


def _draw_two(rng, k, max_chains, dmax):
    """Two species of different chain lengths, up to rho* = dmax."""
    L1, L2 = int(rng.choice([4, 6, 8, 12, 16])), int(rng.choice([4, 6, 8, 12]))
    n1, n2 = int(rng.integers(1, max_chains + 1)), int(rng.integers(1, max_chains // 2 + 2))
    return dict(chains=[L1] * n1 + [L2] * n2, species=[0] * n1 + [1] * n2, density=float(rng.uniform(0.005, dmax)),
                tstar=float(rng.uniform(0.10, 0.60)), canon=bool(rng.random() < 0.6),
                no_hbs=bool(rng.random() < 0.15), no_h=bool(rng.random() < 0.15), gen_seed=70000 + k,
                run_seed=90000 + k)


def _sweep(param_dir, waves, nsys, max_chains, seed, n=10000, dmax=None):
    rng = np.random.default_rng(seed)
    specs = [_draw(rng, k, max_chains) if dmax is None else _draw_two(rng, seed + k, max_chains, dmax)
             for k in range(nsys)]
    systems = [pdmd.generate(s["chains"], density=s["density"], tstar=s["tstar"], seed=s["gen_seed"]) for s in specs]
    eng = pdmd.Engine(waves=waves)
    try:
        ids = [eng.add_system(sysd, seed=s["run_seed"], canon=s["canon"], no_hbs=s["no_hbs"], no_h=s["no_h"],
                              species=s["species"]) for s, sysd in zip(specs, systems)]
        for sid in ids:
            eng.log_enable(sid, n)
        eng.advance(ids, [n] * len(ids))
        for s, sysd, sid in zip(specs, systems, ids):
            st = eng.stats(sid)
            assert st.status == 0, (s, st.status)
            g = eng.log_read(sid, n)
            orc = Oracle(sysd, seed=s["run_seed"], param_dir=param_dir, canon=s["canon"], no_hbs=s["no_hbs"],
                         no_h=s["no_h"], species=s["species"])
            o = orc.run(n, log_cap=n)
            bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                             (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
            assert bad.size == 0, f"{s}: logs diverge at entry {bad[0] + 1}: gpu {g[bad[0]]} oracle {o[bad[0]]}"
            gs, os_ = eng.download(sid), orc.state()
            assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64)), s
            for key in ("identity", "bptnr", "extra", "nptnr", "coltype"):
                assert np.array_equal(gs[key], os_[key]), (s, key)
    finally:
        eng.close()


@pytest.mark.parametrize("waves,nsys,max_chains,seed", [
    (1, 24, 40, 20261201),
    (2, 32, 40, 20261204),
    (4, 64, 40, 20261202),
    # W = 8 fits beside the block minima of systems up to a few thousand beads
    (8, 48, 12, 20261203),
])
def test_sweep(param_dir, waves, nsys, max_chains, seed):
    _sweep(param_dir, waves=waves, nsys=nsys, max_chains=max_chains, seed=seed, n=20000)


@pytest.mark.parametrize("waves,nsys,max_chains,seed,dmax", [
    (4, 48, 12, 20261205, 0.15),
    (8, 40, 8, 20261206, 0.15),
])
def test_sweep_two_species(param_dir, waves, nsys, max_chains, seed, dmax):
    """Two species of different chain lengths per system, up to rho* 0.15."""
    _sweep(param_dir, waves=waves, nsys=nsys, max_chains=max_chains, seed=seed, n=20000, dmax=dmax)
