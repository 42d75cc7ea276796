"""Andersen ghost collisions of the CPU oracle (main.F90:997-1049, initial
draw main.F90:409-415) against the distributions the reference's code draws
from, with the project RNG in place of IFPORT drandm:

* the ghost bead is uniform over 1..N (main.F90:998-1000);
* each resampled velocity component is Maxwell-Boltzmann, Gaussian with
  mean 0 and variance setemp/bm(i) (the polar Box-Muller draws of
  main.F90:1009-1029 scaled by fact = sqrt(-2 setemp bm log(r)/r), then /bm);
* the time to the next ghost is exponential with mean avegtime =
  0.00005/sqrt(setemp) (main.F90:1041-1044, 154-155).

A wrong fact scaling (bm missing, or a factor 2), a wrong avegtime or a
biased bead choice fails these; the thresholds are 4-sigma bounds on the
sample moments and Kolmogorov-Smirnov p > 1e-4.
"""
import math

import numpy as np
import pytest

import pdmd
from oracle import Oracle


def _ks_pvalue(d, n):
    """Asymptotic Kolmogorov distribution: P(D_n > d)."""
    lam = (math.sqrt(n) + 0.12 + 0.11 / math.sqrt(n)) * d
    return max(0.0, min(1.0, 2.0 * sum((-1) ** (k - 1) * math.exp(-2.0 * k * k * lam * lam) for k in range(1, 101))))


def _ks(sample, cdf):
    x = np.sort(np.asarray(sample, np.float64))
    n = x.size
    f = cdf(x)
    d = max(np.max(np.arange(1, n + 1) / n - f), np.max(f - np.arange(0, n) / n))
    return _ks_pvalue(d, n)


@pytest.fixture(scope="module")
def ghosts(param_dir):
    """Every ghost event of a 96-bead system over 4e5 entries: absolute time,
    bead, and the bead's velocity right after the resample."""
    sysd = pdmd.generate([8] * 3, density=0.01, tstar=0.25, seed=77)
    n = 400000
    orc = Oracle(sysd, seed=2024, param_dir=param_dir, canon=True)
    N = orc.N
    log = orc.run(n, log_cap=n)
    # absolute time of each entry: the sync pseudo-event moves the origin
    # by its own time (main.F90:1129-1133)
    sync = log["i"] == N + 2
    origin = np.concatenate([[0.0], np.cumsum(np.where(sync, log["t"], 0.0))[:-1]])
    tabs = origin + log["t"]
    g = np.nonzero(log["i"] == N + 1)[0]
    # velocities right after each ghost: a second run of the same seed
    # stopped after every ghost entry
    orc2 = Oracle(sysd, seed=2024, param_dir=param_dir, canon=True)
    vel, done = [], 0
    for k in g:
        orc2.run(int(k) + 1 - done)
        done = int(k) + 1
        vel.append(orc2.state()["sv"].reshape(N, 6)[log["j"][k] - 1, 3:].copy())
    setemp = 12.0 * 0.25
    bm = np.array([orc.table(5, int(t)) for t in orc.state()["identity"]])
    return dict(N=N, t=tabs[g], bead=log["j"][g], vel=np.array(vel), bm=bm, setemp=setemp)


def test_enough_ghosts(ghosts):
    assert ghosts["t"].size >= 2000


def test_ghost_bead_uniform(ghosts):
    N, b = ghosts["N"], ghosts["bead"]
    assert b.min() >= 1 and b.max() <= N
    cnt = np.bincount(b - 1, minlength=N)
    expect = b.size / N
    chi2 = float(((cnt - expect) ** 2 / expect).sum())
    # chi-square with N-1 dof: mean N-1, sd sqrt(2(N-1)); 5-sigma bound
    assert abs(chi2 - (N - 1)) < 5.0 * math.sqrt(2.0 * (N - 1)), chi2


def test_ghost_velocities_maxwell_boltzmann(ghosts):
    bm = ghosts["bm"][ghosts["bead"] - 1]
    z = (ghosts["vel"] * np.sqrt(bm / ghosts["setemp"])[:, None]).ravel()
    n = z.size
    assert abs(z.mean()) < 4.0 / math.sqrt(n), z.mean()
    assert abs(z.var() - 1.0) < 4.0 * math.sqrt(2.0 / n), z.var()
    p = _ks(z, lambda x: 0.5 * (1.0 + np.vectorize(math.erf)(x / math.sqrt(2.0))))
    assert p > 1e-4, p
    # the variance follows each bead's own mass: per mass class (backbone
    # beads of one mass, side chains of others) the scaled components are
    # unit Gaussians too, so a scaling by the wrong or by no mass fails
    zb = ghosts["vel"] * np.sqrt(bm / ghosts["setemp"])[:, None]
    classes = 0
    for m in np.unique(bm):
        zz = zb[bm == m].ravel()
        if zz.size < 600:
            continue
        classes += 1
        assert abs(zz.var() - 1.0) < 5.0 * math.sqrt(2.0 / zz.size), (m, zz.var())
    assert classes >= 2
    assert np.ptp(bm) > 0.2 * bm.mean()


def test_ghost_interarrival_exponential(ghosts):
    avegtime = 0.00005 / math.sqrt(ghosts["setemp"])
    dt = np.diff(ghosts["t"])
    n = dt.size
    assert abs(dt.mean() / avegtime - 1.0) < 4.0 / math.sqrt(n), dt.mean() / avegtime
    p = _ks(dt / avegtime, lambda x: 1.0 - np.exp(-x))
    assert p > 1e-4, p
