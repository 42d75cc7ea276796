"""-Dchaptype=2 systems on the device against the CPU oracle: the chains
after the first species block are unbonded (no bonds, local cores or
intra-chain wells, no side-chain wells with each other, make_code.f:349-547),
so their beads start scattered 4.5 A from every other bead.  Random systems,
Andersen ghosts on, 2e4 entries at 1 and 4 wavefronts: event logs and final
states bit for bit (the reference-built fixture ref_2species_ch2 pins the
oracle, tests/test_oracle_reference.py, and the device,
tests/test_gpu_reference.py)."""
import numpy as np
import pytest

import pdmd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _system(seed, chains, nfirst):
    sysd = pdmd.generate(chains, density=0.012, tstar=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    sv = np.array(sysd["sv"]).reshape(-1, 6).copy()
    for k in range(4 * sum(chains[:nfirst]), sv.shape[0]):
        while True:
            x = rng.uniform(-0.5, 0.5, 3)
            d = sv[:k, :3] - x
            d -= np.round(d)
            if np.sqrt((d * d).sum(1)).min() * sysd["boxl"] >= 4.5:
                sv[k, :3] = x
                break
    species = np.array([0] * nfirst + [1] * (len(chains) - nfirst), np.int32)
    return dict(sysd, sv=sv.ravel()), species


@pytest.mark.parametrize("waves", [1, 4])
@pytest.mark.parametrize("seed,chains,nfirst", [(61, [12] * 6 + [8] * 4, 6), (62, [16] * 3 + [6] * 10, 3)])
def test_chaptype2_matches_oracle(param_dir, waves, seed, chains, nfirst):
    sysd, species = _system(seed, chains, nfirst)
    n = 20000
    print(f"chaptype2 system {seed}: {4 * sum(chains)} beads at W={waves}", flush=True)
    eng = pdmd.Engine(waves=waves)
    try:
        sid = eng.add_system(sysd, seed=seed, species=species, chaptype2=True)
        eng.log_enable(sid, n)
        eng.advance([sid], [n])
        g = eng.log_read(sid, n)
        gs = eng.download(sid)
        st = eng.stats(sid)
    finally:
        eng.close()
    orc = Oracle(sysd, seed=seed, param_dir=param_dir, species=species, chaptype2=True)
    o = orc.run(n, log_cap=n)
    os_ = orc.state()
    assert st.status == 0
    bad = np.nonzero((g["i"] != o["i"]) | (g["j"] != o["j"]) | (g["type"] != o["type"]) |
                     (g["t"].view(np.int64) != o["t"].view(np.int64)))[0]
    assert bad.size == 0, f"logs differ at entry {bad[0] + 1}: {g[bad[0]]} vs {o[bad[0]]}"
    assert np.array_equal(gs["sv"].view(np.int64), os_["sv"].view(np.int64))
    for k in ("identity", "bptnr", "extra", "nptnr", "coltype"):
        assert np.array_equal(gs[k], os_[k]), k
